// Register-column Jacobi kernels of the batched reference multigrid (FP32;
// used on a level when every group has even nx, ny with ny <= 128 and a
// coarser level of half the size).  Included by aj_batch.hip.
//
// A work item is one x-tile of one instance: W waves, each holding R
// consecutive rows of the instance's whole row width in registers, lane l
// the two columns 2l and 2l + 1.  A Jacobi update reads its x-neighbours
// from the lane's own registers (at a wave's first and last row: from the
// neighbouring wave, whose boundary rows are exchanged through LDS before
// every sweep) and its y-neighbours from the lane's other column and, by a
// DPP wave shift, from the neighbouring lane; the edge-padded wall value
// (the cell itself) is what the shift yields in lane 0 and is selected at
// the last column.  Cells go through the same per-cell functions as the
// LDS tiles (ref_jacobi_val, ref_residual_of, the prolongation cases), so
// both give the same bits.
//
// pre  (PRE): [sums of the squared residual of the input iterate], 2 sweeps,
//      the last-row clamp, then the residual of the result and its 2 x 2
//      restriction: the level's Jacobi launch and its residual-restriction
//      launch in one.  4 halo rows per tile side: two sweeps spoil 2 rows
//      from a tile edge, the residual one more, and restriction pairs start
//      on even rows.
// post:  p_in + prolong(e_c), 2 sweeps, the clamp; 2 halo rows per side.
#pragma once

namespace aj {

template <int W, int R, bool PRE>
struct JacRcGeo {
  static_assert(R % 2 == 0, "even rows per wave");
  static constexpr int HX = PRE ? 4 : 2;  // halo rows per side
  static constexpr int ROWS = W * R;
  static constexpr int U = ROWS - 2 * HX;  // kept rows per tile
  static_assert(U >= 2 && U % 2 == 0, "tile shorter than its halo");
};

// One item's work after its loads.  p, bb: the wave's rows (bb already 0 on
// the last row when bzl).  GEN: the tile reaches row 0 or row nx - 1 (the
// x-edge padding is then applied per row).
template <int W, int R, bool PRE, bool GEN>
__device__ __forceinline__ void jac_rc_item(float (&p)[R][2], const float (&bb)[R][2],
                                            float2 (*xb)[W][2][64], const BatchMgArgs<float>& m,
                                            float* __restrict__ bc, const BatchItem& it, int nx,
                                            int ny, int ncy, int gx0, float ax, float ay,
                                            float inv, double* res_out) {
  using G = JacRcGeo<W, R, PRE>;
  const int wv = wave_id(), lane = (int)(threadIdx.x & 63);
  const int wu = wv > 0 ? wv - 1 : wv, wd = wv < W - 1 ? wv + 1 : wv;
  const int gxw = gx0 + wv * R;
  const int gy = 2 * lane;
  const bool lin = gy < ny;
  const bool hasN1 = gy + 1 < ny - 1;  // column gy + 1 is not the last one
  // edge-padded neighbours of cell (r, s) from the wave's rows q and the
  // exchanged rows above (up) and below (dn)
  auto nbrs = [&](const float (&q)[R][2], float2 up, float2 dn, int r, int s, float* Wv,
                  float* Ev, float* Sv, float* Nv) {
    const float pc = q[r][s];
    *Wv = r > 0 ? q[r - 1][s] : (s ? up.y : up.x);
    *Ev = r < R - 1 ? q[r + 1][s] : (s ? dn.y : dn.x);
    if (GEN) {
      const int gx = gxw + r;
      if (gx == 0) *Wv = pc;
      if (gx == nx - 1) *Ev = pc;
    }
    if (s == 0) {
      *Sv = dpp_from_left(q[r][1], pc);
      *Nv = q[r][1];
    } else {
      *Sv = q[r][0];
      const float n1 = dpp_from_right(q[r][0], pc);
      *Nv = hasN1 ? n1 : pc;
    }
  };
  auto exchange = [&](const float (&q)[R][2], int buf, float2* up, float2* dn) {
    xb[buf][wv][0][lane] = make_float2(q[0][0], q[0][1]);
    xb[buf][wv][1][lane] = make_float2(q[R - 1][0], q[R - 1][1]);
    __syncthreads();
    *up = xb[buf][wu][1][lane];
    *dn = xb[buf][wd][0][lane];
  };
  auto kept = [&](int gx) { return gx >= gx0 + G::HX && gx < gx0 + G::HX + G::U && gx < nx; };
  float2 up, dn;
  exchange(p, 0, &up, &dn);
  if (PRE && res_out) {
    // sum of the squared residual of the input iterate over the kept cells
    double acc = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      if (!kept(gxw + r)) continue;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        float Wv, Ev, Sv, Nv;
        nbrs(p, up, dn, r, s, &Wv, &Ev, &Sv, &Nv);
        const float rv = ref_residual_of(p[r][s], Wv, Ev, Sv, Nv, bb[r][s], ax, ay);
        if (lin) acc += (double)rv * (double)rv;
      }
    }
    const double t = wave_sum(acc);
    if (lane == 0) res_out[wv] = t;
  }
#pragma unroll
  for (int sw = 0; sw < 2; ++sw) {
    if (sw > 0) exchange(p, 1, &up, &dn);
    float np[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        float Wv, Ev, Sv, Nv;
        nbrs(p, up, dn, r, s, &Wv, &Ev, &Sv, &Nv);
        np[r][s] = ref_jacobi_val(Wv, Ev, Sv, Nv, bb[r][s], ax, ay, inv);
      }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      p[r][0] = np[r][0];
      p[r][1] = np[r][1];
    }
  }
  if (m.clamp) {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (gxw + r == nx - 1) {
        p[r][0] = 0.0f;
        p[r][1] = 0.0f;
      }
  }
  float* out = m.p_out + it.off;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (kept(gxw + r) && lin)
      *reinterpret_cast<float2*>(out + (long)(gxw + r) * ny + gy) = make_float2(p[r][0], p[r][1]);
  if (PRE) {
    // residual of the smoothed (clamped) iterate and its restriction
    // (multigrid_solver.py:108-132, 198-199)
    exchange(p, 0, &up, &dn);
    float* bcl = bc + it.coff;
#pragma unroll
    for (int r = 0; r < R; r += 2) {
      if (!kept(gxw + r)) continue;  // pairs: gxw + r even, both rows kept or neither
      float rr[2][2];
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
          float Wv, Ev, Sv, Nv;
          nbrs(p, up, dn, r + h, s, &Wv, &Ev, &Sv, &Nv);
          rr[h][s] = ref_residual_of(p[r + h][s], Wv, Ev, Sv, Nv, bb[r + h][s], ax, ay);
        }
      if (lin)
        bcl[(long)((gxw + r) >> 1) * ncy + lane] =
            0.25f * (((rr[0][0] + rr[1][0]) + rr[0][1]) + rr[1][1]);
    }
  }
}

template <int W, int R, bool PRE>
__global__ __launch_bounds__(W * 64) void kb_jacobi_rc(BatchWork w, BatchMgArgs<float> m,
                                                       float* __restrict__ bc) {
  using G = JacRcGeo<W, R, PRE>;
  constexpr int NQ = R / 2 + 1;  // coarse rows of a wave's rows (post)
  __shared__ float2 xb[2][W][2][64];
  const int wv = wave_id(), lane = (int)(threadIdx.x & 63);
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const BatchItem it = batch_item(w, q.base + q.t);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    // fine-grid spacing on every level (multigrid_solver.py:158-171)
    const float ax = 1.0f / ((float)w.dx[g] * (float)w.dx[g]);
    const float ay = 1.0f / ((float)w.dy[g] * (float)w.dy[g]);
    const float inv = 1.0f / (2.0f * (ax + ay));
    const int gx0 = it.sub * G::U - G::HX;  // even
    const int gxw = gx0 + wv * R;
    const int gy = 2 * lane;
    const bool lin = gy < ny;
    const int gyl = lin ? gy : 0;  // lanes past the row read column 0
    float p[R][2], bb[R][2];
    float cr[NQ];
    if (!PRE) {
      const int ncx = w.cnx[g], ncy = w.cny[g];
      const float* ec = m.e_c + it.coff;
      const int Jc = min(lane, ncy - 1);
#pragma unroll
      for (int k = 0; k < NQ; ++k) cr[k] = ec[(long)clampi((gxw >> 1) + k, 0, ncx - 1) * ncy + Jc];
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gx = clampi(gxw + r, 0, nx - 1);
      const long gi = it.off + (long)gx * ny + gyl;
      const float2 vb = *reinterpret_cast<const float2*>(m.b + gi);
      const bool zb = m.bzl && gx == nx - 1;
      bb[r][0] = (lin && !zb) ? vb.x : 0.0f;
      bb[r][1] = (lin && !zb) ? vb.y : 0.0f;
      const float2 vp =
          m.p_in ? *reinterpret_cast<const float2*>(m.p_in + gi) : make_float2(0.0f, 0.0f);
      p[r][0] = lin ? vp.x : 0.0f;
      p[r][1] = lin ? vp.y : 0.0f;
    }
    if (gated(m.stats + it.inst)) continue;
    if (!PRE) {
      // start value p_in + prolong(e_c) (multigrid_solver.py:134-156): lane
      // l holds coarse column l, its right neighbour comes from lane l + 1
      const int ncx = w.cnx[g], ncy = w.cny[g];
      const bool jlast = lane == ncy - 1;  // column 2l + 1 is the last fine column
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int k = r >> 1;
        const bool io = r & 1;  // gxw is even
        const bool zrow = gxw + r == 2 * ncx - 1;
        const float c00 = cr[k], c10 = cr[k + 1];
        const float c01 = dpp_from_right(c00, 0.0f), c11 = dpp_from_right(c10, 0.0f);
        const float pr0 = io ? 0.5f * (c00 + c10) : c00;
        const float s1 = c00 + (io ? c10 : c01);
        const float pr1 = io ? 0.25f * ((s1 + c01) + c11) : 0.5f * s1;
        p[r][0] = p[r][0] + (zrow ? 0.0f : pr0);
        p[r][1] = p[r][1] + ((zrow || jlast) ? 0.0f : pr1);
      }
    }
    double* res_out = (PRE && m.res_partial)
                          ? m.res_partial + (long)it.inst * m.res_stride + W * it.sub
                          : nullptr;
    if (gx0 >= 1 && gx0 + G::ROWS <= nx - 1)
      jac_rc_item<W, R, PRE, false>(p, bb, xb, m, bc, it, nx, ny, ny >> 1, gx0, ax, ay, inv,
                                    res_out);
    else
      jac_rc_item<W, R, PRE, true>(p, bb, xb, m, bc, it, nx, ny, ny >> 1, gx0, ax, ay, inv,
                                   res_out);
    __syncthreads();  // the exchange buffers are reused by the next item
  }
}

}  // namespace aj
