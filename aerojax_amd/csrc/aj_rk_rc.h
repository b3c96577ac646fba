// Register-column RK predictor (FP32, no hyperviscosity): further launch
// candidates beside the LDS tiles of rk_tile, evaluating every cell with
// the same per-cell functions (rk_rhs_cell, rk_stage, rk_inv_d, rk_div), so
// both give the same bits.  Used by the single-grid entry points (tuned
// candidates) and the batched driver (rows of at most 128 cells).
//
// A tile is W waves x R rows x 128 columns: wave w holds rows
// gx0 + w R .. + R - 1 of u, v and their per-cell constants in registers,
// lane l the columns gy0 + 2l and gy0 + 2l + 1.  A stage reads its
// x-neighbours from the lane's own registers (at a wave's first and last
// row: the neighbouring wave's, exchanged through LDS before every stage)
// and its y-neighbours from the lane's other column and, by a DPP wave
// shift, the neighbouring lane's; edge padding (a neighbour outside the
// domain is the cell itself) is a per-row choice for the inlet and outlet
// rows and a per-lane one at the walls.
//
// Temporal blocking: the stages spoil one more row and column per stage
// from the tile edges (one more with the rhs epilogue), so a tile keeps
// rows [gx0 + HX, gx0 + ROWS - HX) and, in strips narrower than the grid,
// columns [gy0 + 4, gy0 + 124); a grid of at most 128 (even) columns is one
// strip with gy0 = 0 and keeps every column.
//
// Writes follow rk_tile: a kept cell of the inlet row writes the inflow,
// an interior kept cell its masked value and the boundary cells it is the
// source of (the outlet row copies row nx - 2, a wall takes the slip or
// no-slip value of its neighbour, v = 0); every kept cell writes its rhs
// (0 off the interior).  Both columns of a wall's source and target lie in
// one lane (ny even), so every store is a float2 of the lane's two cells.
#pragma once

namespace aj {

template <int W, int R, bool FAST, bool DIV>
struct RkRcGeo {
  static constexpr int S = FAST ? 2 : 3;
  static constexpr int HX = S + (DIV ? 1 : 0);  // halo rows per tile side
  static constexpr int HY = 4;                  // halo columns per strip side (strips)
  static constexpr int ROWS = W * R;
  static constexpr int U = ROWS - 2 * HX;       // kept rows (interior tiles)
  static constexpr int UY = 128 - 2 * HY;       // kept columns of a strip
  static_assert(U >= 2, "tile shorter than its halo");
  // x-tiles with no halo on the inlet and outlet sides (edge_x_tiles)
  __host__ __device__ static int x_tiles(int n) { return edge_x_tiles(n, ROWS, HX); }
  __host__ __device__ static void x_tile(int xt, int n, int* gx0, int* k0, int* k1) {
    edge_x_tile(xt, n, ROWS, HX, gx0, k0, k1);
  }
};

// Per-launch scalars of the register-column predictor.
struct RkRcArgs {
  const float *u, *v, *mask, *sdf, *nu_sgs, *nu_field;
  float *u_out, *v_out, *rhs;
  int nx, ny;
  float dt, nu, inv_dx, inv_dy, inv_dx2, inv_dy2, U_inf, inv_eta, sdf_scale;
  float inv_2dx, inv_2dy, inv_dt;
  int slip, rhs_masked;
};

// The whole tile: loads, stages, final mask, boundary values, rhs, stores.
// gx0, gy0: origin; kx0, kx1: kept rows; ky0, ky1: kept columns (even).
// GEN: the tile reaches rows 0, 1, nx - 2 or nx - 1 (per-row choices then
// apply).
template <int W, int R, bool FAST, bool DIV, bool GEN>
__device__ __forceinline__ void rk_rc_body(const RkRcArgs& a, int gx0, int kx0, int kx1, int gy0,
                                           int ky0, int ky1,
                                           float4 (*xb)[W][2][64], const float (&u0)[R][2],
                                           const float (&v0)[R][2], const float (&iD)[R][2],
                                           const float (&nt)[R][2], const float (&mk)[R][2]) {
  using G = RkRcGeo<W, R, FAST, DIV>;
  const int wv = wave_id(), lane = (int)(threadIdx.x & 63);
  const int wu = wv > 0 ? wv - 1 : wv, wd = wv < W - 1 ? wv + 1 : wv;
  const int nx = a.nx, ny = a.ny;
  const int gxw = gx0 + wv * R;
  const int gy = gy0 + 2 * lane;
  const bool hasS0 = gy > 0, hasN1 = gy + 1 < ny - 1;
  float cu[R][2], cv[R][2], k1u[R][2], k1v[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      cu[r][s] = u0[r][s];
      cv[r][s] = v0[r][s];
      k1u[r][s] = 0.0f;
      k1v[r][s] = 0.0f;
    }
  auto exchange = [&](const float (&qu)[R][2], const float (&qv)[R][2], int buf, float4* up,
                      float4* dn) {
    xb[buf][wv][0][lane] = make_float4(qu[0][0], qu[0][1], qv[0][0], qv[0][1]);
    xb[buf][wv][1][lane] = make_float4(qu[R - 1][0], qu[R - 1][1], qv[R - 1][0], qv[R - 1][1]);
    __syncthreads();
    *up = xb[buf][wu][1][lane];
    *dn = xb[buf][wd][0][lane];
  };
  // edge-padded neighbours of cell (r, s) of field q (x: up/dn hold the
  // rows above and below the wave; comp 0 = u, 1 = v)
  auto nbrs = [&](const float (&q)[R][2], float4 up, float4 dn, int comp, int r, int s,
                  float* Wv, float* Ev, float* Sv, float* Nv) {
    const float c = q[r][s];
    const float upv = comp == 0 ? (s ? up.y : up.x) : (s ? up.w : up.z);
    const float dnv = comp == 0 ? (s ? dn.y : dn.x) : (s ? dn.w : dn.z);
    *Wv = r > 0 ? q[r - 1][s] : upv;
    *Ev = r < R - 1 ? q[r + 1][s] : dnv;
    if (GEN) {
      const int gx = gxw + r;
      if (gx == 0) *Wv = c;
      if (gx == nx - 1) *Ev = c;
    }
    if (s == 0) {
      const float l = dpp_from_left(q[r][1], c);
      *Sv = hasS0 ? l : c;
      *Nv = q[r][1];
    } else {
      const float rt = dpp_from_right(q[r][0], c);
      *Sv = q[r][0];
      *Nv = hasN1 ? rt : c;
    }
  };
  float4 up, dn;
#pragma unroll
  for (int st = 1; st <= G::S; ++st) {
    exchange(cu, cv, (st - 1) & 1, &up, &dn);
    float nu_[R][2], nv_[R][2];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        float uW, uE, uS, uN, vW, vE, vS, vN;
        nbrs(cu, up, dn, 0, r, s, &uW, &uE, &uS, &uN);
        nbrs(cv, up, dn, 1, r, s, &vW, &vE, &vS, &vN);
        float ku, kv;
        rk_rhs_cell(cu[r][s], uW, uE, uS, uN, cv[r][s], vW, vE, vS, vN, nt[r][s], a.inv_dx,
                    a.inv_dy, a.inv_dx2, a.inv_dy2, ku, kv);
        if (FAST && st == 1) {
          k1u[r][s] = ku;
          k1v[r][s] = kv;
        }
        nu_[r][s] = rk_stage<float, FAST>(st, u0[r][s], ku, k1u[r][s], a.dt, iD[r][s]);
        nv_[r][s] = rk_stage<float, FAST>(st, v0[r][s], kv, k1v[r][s], a.dt, iD[r][s]);
      }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        cu[r][s] = nu_[r][s];
        cv[r][s] = nv_[r][s];
      }
  }
  // final values, masked (rk3_simple_new.py:332-336)
  float fu[R][2], fv[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      fu[r][s] = masked(cu[r][s], mk[r][s]);
      fv[r][s] = masked(cv[r][s], mk[r][s]);
    }
  float rh[R][2];
  if (DIV) {
    // divergence of the final values with the values the BCs give the
    // neighbours: inlet u = U_inf, outlet row = copy of row nx - 2, wall v = 0
    exchange(fu, fv, G::S & 1, &up, &dn);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int gx = gxw + r;
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        float uW, uE, vS, vN;
        uW = r > 0 ? fu[r - 1][s] : (s ? up.y : up.x);
        uE = r < R - 1 ? fu[r + 1][s] : (s ? dn.y : dn.x);
        if (GEN) {
          if (gx == nx - 2) uE = fu[r][s];
          if (gx == 1) uW = a.U_inf;
        }
        if (s == 0) {
          vS = dpp_from_left(fv[r][1], 0.0f);
          vN = (gy == ny - 2) ? 0.0f : fv[r][1];
        } else {
          vS = (gy + 1 == 1) ? 0.0f : fv[r][0];
          vN = dpp_from_right(fv[r][0], 0.0f);
        }
        const float d = rk_div(uE, uW, vN, vS, a.inv_2dx, a.inv_2dy);
        const int col = gy + s;
        const bool inter = gx > 0 && gx < nx - 1 && col > 0 && col < ny - 1;
        rh[r][s] = inter ? (a.rhs_masked ? d * mk[r][s] * a.inv_dt : d * a.inv_dt) : 0.0f;
      }
    }
  }
  // stores of the kept cells (rk3_simple_new.py:264-279 for the boundary values)
  // both columns of a lane are inside the grid or neither (gy even, ny even)
  const bool lkeep = gy >= ky0 && gy < ky1 && gy >= 0 && gy < ny;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gx = gxw + r;
    const bool rkeep = gx >= kx0 && gx < kx1;
    if (!(rkeep && lkeep)) continue;
    const long g = (long)gx * ny + gy;
    if (DIV) *reinterpret_cast<float2*>(a.rhs + g) = make_float2(rh[r][0], rh[r][1]);
    if (gx == nx - 1) continue;  // written by its source row nx - 2
    float2 ou, ov;
    if (gx == 0) {
      const bool w0 = gy == 0, w1 = gy + 1 == ny - 1;
      ou = make_float2((w0 && !a.slip) ? 0.0f : a.U_inf, (w1 && !a.slip) ? 0.0f : a.U_inf);
      ov = make_float2(0.0f, 0.0f);
    } else {
      ou = make_float2(fu[r][0], fu[r][1]);
      ov = make_float2(fv[r][0], fv[r][1]);
      if (gy == 0) {  // wall column 0 from its source column 1
        ou.x = a.slip ? fu[r][1] : 0.0f;
        ov.x = 0.0f;
      }
      if (gy + 1 == ny - 1) {  // wall column ny - 1 from column ny - 2
        ou.y = a.slip ? fu[r][0] : 0.0f;
        ov.y = 0.0f;
      }
    }
    *reinterpret_cast<float2*>(a.u_out + g) = ou;
    *reinterpret_cast<float2*>(a.v_out + g) = ov;
    if (gx == nx - 2) {  // the outlet row copies row nx - 2
      *reinterpret_cast<float2*>(a.u_out + g + ny) = ou;
      *reinterpret_cast<float2*>(a.v_out + g + ny) = ov;
    }
  }
}

// Loads of a tile and the body.  kx0, kx1: kept rows; ky0, ky1: kept
// columns.
template <int W, int R, bool FAST, bool DIV>
__device__ __forceinline__ void rk_rc_tile(const RkRcArgs& a, int gx0, int kx0, int kx1, int gy0,
                                           int ky0, int ky1, float4 (*xb)[W][2][64]) {
  const int wv = wave_id(), lane = (int)(threadIdx.x & 63);
  const int nx = a.nx, ny = a.ny;
  const int gxw = gx0 + wv * R;
  const int gyl = clampi(gy0 + 2 * lane, 0, ny - 2);  // lanes outside the grid read an edge pair
  float u0[R][2], v0[R][2], iD[R][2], nt[R][2], mk[R][2];
  {
    float2 ru[R], rv[R], rc[R], rm[R], rs[R], rf[R];
    const float* csrc = a.sdf ? a.sdf : a.mask;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const long g = (long)clampi(gxw + r, 0, nx - 1) * ny + gyl;
      ru[r] = *reinterpret_cast<const float2*>(a.u + g);
      rv[r] = *reinterpret_cast<const float2*>(a.v + g);
      rm[r] = *reinterpret_cast<const float2*>(a.mask + g);
      rc[r] = a.sdf ? *reinterpret_cast<const float2*>(csrc + g) : rm[r];
      rs[r] = a.nu_sgs ? *reinterpret_cast<const float2*>(a.nu_sgs + g) : make_float2(0.f, 0.f);
      rf[r] = a.nu_field ? *reinterpret_cast<const float2*>(a.nu_field + g)
                         : make_float2(a.nu, a.nu);
    }
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        u0[r][s] = s ? ru[r].y : ru[r].x;
        v0[r][s] = s ? rv[r].y : rv[r].x;
        mk[r][s] = s ? rm[r].y : rm[r].x;
        const float c = s ? rc[r].y : rc[r].x;
        iD[r][s] = rk_inv_d(rk_chi(c, a.sdf != nullptr, a.sdf_scale), a.dt, a.inv_eta);
        // nu_total = nu (+ nu_sgs), nu a scalar or a field (rk3_simple_new.py:304-309)
        const float nf = s ? rf[r].y : rf[r].x, ns = s ? rs[r].y : rs[r].x;
        nt[r][s] = a.nu_sgs ? nf + ns : nf;
      }
  }
  using G = RkRcGeo<W, R, FAST, DIV>;
  if (gx0 >= 2 && gx0 + G::ROWS <= nx - 2)
    rk_rc_body<W, R, FAST, DIV, false>(a, gx0, kx0, kx1, gy0, ky0, ky1, xb, u0, v0, iD, nt, mk);
  else
    rk_rc_body<W, R, FAST, DIV, true>(a, gx0, kx0, kx1, gy0, ky0, ky1, xb, u0, v0, iD, nt, mk);
}

// Single grid: grid (8 * ceil(XT * YT / 8)); tiles numbered so that each XCD
// takes a contiguous range (YT strips fastest).  A grid of at most 128
// columns is one strip keeping every column.
template <int W, int R, bool FAST, bool DIV>
__global__ __launch_bounds__(W * 64) void k_rk_rc(RkRcArgs a, int XT, int YT) {
  using G = RkRcGeo<W, R, FAST, DIV>;
  __shared__ float4 xb[2][W][2][64];
  const int T = XT * YT, per = (int)(gridDim.x >> 3), k = (int)blockIdx.x;
  const int t = (k & 7) * per + (k >> 3);
  if (t >= T) return;
  const int xt = t / YT, yt = t - xt * YT;
  int gx0, kx0, kx1;
  G::x_tile(xt, a.nx, &gx0, &kx0, &kx1);
  int gy0, ky0, ky1;
  if (a.ny <= 128) {
    gy0 = 0;
    ky0 = 0;
    ky1 = a.ny;
  } else {
    gy0 = yt * G::UY - G::HY;
    ky0 = gy0 + G::HY;
    ky1 = ky0 + G::UY;
  }
  rk_rc_tile<W, R, FAST, DIV>(a, gx0, kx0, kx1, gy0, ky0, ky1, xb);
}

}  // namespace aj
