// Device side of the step kernels (RK predictor tile, per-cell pressure rhs
// and projection), shared by the single-grid entry points and the batched
// driver.  Layout: (nx, ny) row-major, y contiguous.
#pragma once

#include "aj_common.h"

namespace aj {

// --------------------------------------------------------------------------
// Smagorinsky eddy viscosity at (i, j) (solver/les_models.py:22-54,
// 179-200): (Cs delta)^2 sqrt(2 (Sxx^2 + Syy^2 + 2 Sxy^2) + 1e-10) with
// periodic jnp.roll neighbours (solver/operators.py:11-14, 35-38);
// coef = (Cs delta)^2
// --------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T les_cell(const T* __restrict__ u, const T* __restrict__ v, int nx,
                                      int ny, int i, int j, T inv_2dx, T inv_2dy, T coef) {
  const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
  const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
  const long c = (long)i * ny;
  const T du_dx = (u[(long)ip * ny + j] - u[(long)im * ny + j]) * inv_2dx;
  const T du_dy = (u[c + jp] - u[c + jm]) * inv_2dy;
  const T dv_dx = (v[(long)ip * ny + j] - v[(long)im * ny + j]) * inv_2dx;
  const T dv_dy = (v[c + jp] - v[c + jm]) * inv_2dy;
  const T sxy = T(0.5) * (du_dy + dv_dx);
  const T smag = sqrt(T(2) * (du_dx * du_dx + dv_dy * dv_dy + T(2) * sxy * sxy) + T(1e-10));
  return coef * smag;
}

// --------------------------------------------------------------------------
// RK predictor (advection_schemes/rk3_simple_new.py:175-358)
// --------------------------------------------------------------------------

template <typename T>
struct RkArgs {
  const T* u;
  const T* v;
  const T* mask;
  const T* sdf;     // may be null -> chi = 1 - mask
  const T* nu_sgs;  // may be null
  const T* nu_field;  // may be null: per-cell viscosity replacing nu
  T* u_out;
  T* v_out;
  T* rhs;  // fused pressure rhs (DIV tiles): div(u*, v*) [* mask] / dt
  int nx, ny;
  T dt, nu, inv_dx, inv_dy, inv_dx2, inv_dy2, U_inf, inv_eta, sdf_scale;
  T hyper;  // nu_hyper_ratio * dx * dy (the biharmonic coefficient is nu * hyper)
  T inv_2dx, inv_2dy, inv_dt;  // rhs coefficients (DIV tiles)
  int slip;
  int rhs_masked;  // DIV tiles: multiply the divergence by the mask
};

// The per-cell arithmetic of the RK predictor, written out operation by
// operation (explicit fused multiply-adds): the LDS tiles and the
// register-column kernels evaluate every cell with these functions, so any
// of them gives the same bits.
//
// upwind advection vel * df/dh (rk3_simple_new.py:287-298): backward
// difference where vel > 0, forward otherwise
template <typename T>
__device__ __forceinline__ T rk_upwind(T vel, T fc, T fm, T fp, T inv_h) {
#pragma clang fp contract(off)
  const T d = (vel > T(0)) ? fc - fm : fp - fc;
  return vel * d * inv_h;
}

// 5-point laplacian from the edge-padded neighbours (rk3_simple_new.py:215-219)
template <typename T>
__device__ __forceinline__ T rk_lap(T c, T W, T E, T S, T N, T inv_dx2, T inv_dy2) {
#pragma clang fp contract(off)
  return fma(fma(T(-2), c, E + W), inv_dx2, fma(T(-2), c, N + S) * inv_dy2);
}

// rhs = -upwind(f; u, v) + nu_tot lap(f) for f = u and f = v
template <typename T>
__device__ __forceinline__ void rk_rhs_cell(T uc, T uW, T uE, T uS, T uN, T vc, T vW, T vE, T vS,
                                            T vN, T nu_tot, T inv_dx, T inv_dy, T inv_dx2,
                                            T inv_dy2, T& ru, T& rv) {
#pragma clang fp contract(off)
  const T adv_u = rk_upwind(uc, uc, uW, uE, inv_dx) + rk_upwind(vc, uc, uS, uN, inv_dy);
  const T adv_v = rk_upwind(uc, vc, vW, vE, inv_dx) + rk_upwind(vc, vc, vS, vN, inv_dy);
  ru = fma(nu_tot, rk_lap(uc, uW, uE, uS, uN, inv_dx2, inv_dy2), -adv_u);
  rv = fma(nu_tot, rk_lap(vc, vW, vE, vS, vN, inv_dx2, inv_dy2), -adv_v);
}

// Brinkman penalisation chi: sigmoid(-sdf / (0.1 dx)) from the SDF
// (rk3_simple_new.py:198; scale = 1 / (0.1 dx)), else 1 - mask
template <typename T>
__device__ __forceinline__ T rk_chi(T c, bool from_sdf, T scale) {
#pragma clang fp contract(off)
  return from_sdf ? T(1) / (T(1) + exp(c * scale)) : T(1) - c;
}

// 1 / (1 + dt chi / eta): the Brinkman division of every stage
// (rk3_simple_new.py:196-213)
template <typename T>
__device__ __forceinline__ T rk_inv_d(T chi, T dt, T inv_eta) {
#pragma clang fp contract(off)
  return T(1) / fma(dt * chi, inv_eta, T(1));
}

// stage s of RK3 (rk3_simple_new.py:315-331) or RK2 (FAST, :339-347) from
// the step's start value f0, this stage's rhs k (RK2: k1 the first stage's)
template <typename T, bool FAST>
__device__ __forceinline__ T rk_stage(int s, T f0, T k, T k1, T dt, T iD) {
#pragma clang fp contract(off)
  if (s == 1) return fma(dt, k, f0) * iD;
  if (FAST) return fma(T(0.5) * dt, k1 + k, f0) * iD;
  const T fs = fma(dt, k, f0);
  if (s == 2) return fma(T(0.75), f0, T(0.25) * fs) * iD;
  return fma(T(1.0 / 3.0), f0, T(2.0 / 3.0) * fs) * iD;
}

// pressure rhs from the central differences of u* along x and v* along y
// (divergence_nonperiodic, solver/operators.py:81-96), before the mask and 1/dt
template <typename T>
__device__ __forceinline__ T rk_div(T uE, T uW, T vN, T vS, T inv_2dx, T inv_2dy) {
#pragma clang fp contract(off)
  return fma(uE - uW, inv_2dx, (vN - vS) * inv_2dy);
}

// One stencil evaluation of rhs = -upwind(f; u, v) + nu_tot * lap(f)
// [- nu_h * lap(lap f)] for f = u and f = v at LDS cell l.
template <typename T, bool HYPER>
__device__ __forceinline__ void rk_rhs(const T* fu, const T* fv, const T* lu, const T* lv,
                                       int l, int lW, int lE, int lS, int lN, T nu_tot,
                                       T nu_h, const RkArgs<T>& a, T& ru, T& rv) {
  rk_rhs_cell(fu[l], fu[lW], fu[lE], fu[lS], fu[lN], fv[l], fv[lW], fv[lE], fv[lS], fv[lN],
              nu_tot, a.inv_dx, a.inv_dy, a.inv_dx2, a.inv_dy2, ru, rv);
  if constexpr (HYPER) {
    const T bu = rk_lap(lu[l], lu[lW], lu[lE], lu[lS], lu[lN], a.inv_dx2, a.inv_dy2);
    const T bv = rk_lap(lv[l], lv[lW], lv[lE], lv[lS], lv[lN], a.inv_dx2, a.inv_dy2);
    ru = fma(-nu_h, bu, ru);
    rv = fma(-nu_h, bv, rv);
  }
}

template <typename T>
__device__ __forceinline__ T masked(T val, T m) {
#pragma clang fp contract(off)
  return (m > T(0.5)) ? val : T(0);
}

// LDS elements of a fused RK tile
template <int TX, int TY, bool FAST, bool HYPER, bool DIV>
struct RkTileShape {
  static constexpr int S = FAST ? 2 : 3;
  static constexpr int R1 = HYPER ? 2 : 1;          // per-stage stencil radius
  static constexpr int H = S * R1 + (DIV ? 1 : 0);  // total halo
  static constexpr int RN = (TX + 2 * H) * (TY + 2 * H);
  static constexpr int RF = DIV ? (TX + 2) * (TY + 2) : 0;  // final values on margin 1
  static constexpr int N = RN * (HYPER ? 11 : 8) + (DIV ? RN + 2 * RF : 0);
};

// Fused RK predictor on the tile with origin (x0, y0) (LDS at `sm`,
// RkTileShape::N elements).  FAST = RK2 (2 stages), else RK3.  HYPER adds
// the biharmonic term (per-stage stencil radius 2).  DIV also writes the
// pressure rhs of step_pure (divergence_nonperiodic(u*, v*) * mask / dt,
// solver/operators.py:81-96, step.py:194-199): the final stage runs on the
// tile plus one ring (one more halo cell), and the divergence reads those
// values with the boundary values the BCs give them.
template <typename T, int TX, int TY, bool FAST, bool HYPER, int NTH, bool DIV = false>
__device__ __forceinline__ void rk_tile(const RkArgs<T>& a, int x0, int y0, T* sm) {
  using Sh = RkTileShape<TX, TY, FAST, HYPER, DIV>;
  constexpr int S = Sh::S;
  constexpr int R1 = Sh::R1;         // per-stage stencil radius
  constexpr int H = Sh::H;           // total halo
  constexpr int RH = TX + 2 * H;     // region rows (x)
  constexpr int RW = TY + 2 * H;     // region cols (y, contiguous)
  constexpr int RN = RH * RW;

  T* s_u = sm;              // u (margin H)
  T* s_v = s_u + RN;
  T* s_iD = s_v + RN;       // 1/(1 + dt chi/eta) (margin H - R1)
  T* s_nt = s_iD + RN;      // nu + nu_sgs        (margin H - R1)
  T* s_au = s_nt + RN;      // stage buffers
  T* s_av = s_au + RN;
  T* s_bu = s_av + RN;
  T* s_bv = s_bu + RN;      // RK3: stage-2 output; RK2: k1
  T* s_lu = s_bv + RN;      // HYPER only: laplacians of the stage input
  T* s_lv = s_lu + RN;
  T* s_nh = s_lv + RN;      // HYPER only: biharmonic coefficient nu * hyper
  T* s_m = sm + RN * (HYPER ? 11 : 8);  // DIV only: mask
  T* s_fu = s_m + RN;                   // DIV only: final u*, v* (masked) on margin 1
  T* s_fv = s_fu + Sh::RF;

  const int nx = a.nx, ny = a.ny;
  const int gx0 = x0 - H, gy0 = y0 - H;  // global index of LDS (0, 0)
  const int tid = threadIdx.x;
  const bool has_nsgs = a.nu_sgs != nullptr;
  const bool has_nuf = a.nu_field != nullptr;
  const bool has_sdf = a.sdf != nullptr;

  // mask of this thread's tile cells (final stage), loaded up front
  constexpr int NT = (TX * TY + NTH - 1) / NTH;
  T rmask[NT];
#pragma unroll
  for (int it = 0; it < NT; ++it) {
    const int idx = min(tid + it * NTH, TX * TY - 1);
    const int gx = clampi(x0 + idx / TY, 0, nx - 1), gy = clampi(y0 + idx % TY, 0, ny - 1);
    rmask[it] = a.mask[(long)gx * ny + gy];
  }
  // ---- load u, v on margin H; chi factor and nu_tot on margin H - R1 ----
  // All loads of a thread are issued before any is consumed (addresses are
  // clamped into the domain, so every load is in bounds; slots outside the
  // domain hold edge copies and are never read by the stencils).
  {
    constexpr int NIT = (RN + NTH - 1) / NTH;
    T ru[NIT], rv[NIT], rc[NIT], rn[NIT], rf[NIT], rm[DIV ? NIT : 1];
    const T* csrc = has_sdf ? a.sdf : a.mask;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int idx = min(tid + it * NTH, RN - 1);
      const int lx = idx / RW, ly = idx - lx * RW;
      const long g = (long)clampi(gx0 + lx, 0, nx - 1) * ny + clampi(gy0 + ly, 0, ny - 1);
      ru[it] = a.u[g];
      rv[it] = a.v[g];
      rc[it] = csrc[g];
      rn[it] = has_nsgs ? a.nu_sgs[g] : T(0);
      rf[it] = has_nuf ? a.nu_field[g] : a.nu;
      if constexpr (DIV) rm[it] = a.mask[g];
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int idx = tid + it * NTH;
      if (idx >= RN) break;
      s_u[idx] = ru[it];
      s_v[idx] = rv[it];
      s_iD[idx] = rk_inv_d(rk_chi(rc[it], has_sdf, a.sdf_scale), a.dt, a.inv_eta);
      // nu_total = nu (+ nu_sgs) with nu a scalar or a field
      // (rk3_simple_new.py:304-309)
      s_nt[idx] = has_nsgs ? rf[it] + rn[it] : rf[it];
      if constexpr (HYPER) s_nh[idx] = rf[it] * a.hyper;
      if constexpr (DIV) s_m[idx] = rm[it];
    }
  }
  __syncthreads();

  const T* fu = s_u;
  const T* fv = s_v;
#pragma unroll
  for (int s = 1; s <= S; ++s) {
    const int m = (S - s) * R1 + (DIV ? 1 : 0);  // output margin of this stage
    if constexpr (HYPER) {
      // laplacian of the stage input on margin m + 1
      const int mm = m + 1;
      const int W = TY + 2 * mm, Hh = TX + 2 * mm;
      for (int idx = tid; idx < W * Hh; idx += NTH) {
        const int lx = H - mm + idx / W, ly = H - mm + idx % W;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
        const int l = lx * RW + ly;
        const int lW = gx > 0 ? l - RW : l, lE = gx < nx - 1 ? l + RW : l;
        const int lS = gy > 0 ? l - 1 : l, lN = gy < ny - 1 ? l + 1 : l;
        s_lu[l] = rk_lap(fu[l], fu[lW], fu[lE], fu[lS], fu[lN], a.inv_dx2, a.inv_dy2);
        s_lv[l] = rk_lap(fv[l], fv[lW], fv[lE], fv[lS], fv[lN], a.inv_dx2, a.inv_dy2);
      }
      __syncthreads();
    }
    T* ou = (s == 1) ? s_au : s_bu;
    T* ov = (s == 1) ? s_av : s_bv;
    // stage value at LDS cell l (RK3: rk3_simple_new.py:315-331; RK2: :339-347)
    auto stage = [&](int l, int gx, int gy, T& nu_, T& nv_) {
      const int lW = gx > 0 ? l - RW : l, lE = gx < nx - 1 ? l + RW : l;
      const int lS = gy > 0 ? l - 1 : l, lN = gy < ny - 1 ? l + 1 : l;
      T ku, kv;
      rk_rhs<T, HYPER>(fu, fv, s_lu, s_lv, l, lW, lE, lS, lN, s_nt[l], HYPER ? s_nh[l] : T(0), a,
                       ku, kv);
      const T u0 = s_u[l], v0 = s_v[l], iD = s_iD[l];
      T k1u = T(0), k1v = T(0);
      if constexpr (FAST) {
        if (s == 1) {
          // k1 is needed again at the final stage (margin 0 within margin R1)
          s_bu[l] = ku;
          s_bv[l] = kv;
        } else {
          k1u = s_bu[l];
          k1v = s_bv[l];
        }
      }
      nu_ = rk_stage<T, FAST>(s, u0, ku, k1u, a.dt, iD);
      nv_ = rk_stage<T, FAST>(s, v0, kv, k1v, a.dt, iD);
    };
    if (s < S) {
      const int W = TY + 2 * m, Hh = TX + 2 * m;
      for (int idx = tid; idx < W * Hh; idx += NTH) {
        const int lx = H - m + idx / W, ly = H - m + idx % W;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
        const int l = lx * RW + ly;
        T nu_, nv_;
        stage(l, gx, gy, nu_, nv_);
        ou[l] = nu_;
        ov[l] = nv_;
      }
      __syncthreads();
      fu = ou;
      fv = ov;
      continue;
    }
    if constexpr (DIV) {
      // final stage, masked, on the tile plus one ring
      constexpr int FW = TY + 2, FH = TX + 2;
      for (int idx = tid; idx < FW * FH; idx += NTH) {
        const int lx = H - 1 + idx / FW, ly = H - 1 + idx % FW;
        const int gx = gx0 + lx, gy = gy0 + ly;
        if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
        const int l = lx * RW + ly;
        T nu_, nv_;
        stage(l, gx, gy, nu_, nv_);
        const T m0 = s_m[l];
        s_fu[idx] = masked(nu_, m0);
        s_fv[idx] = masked(nv_, m0);
      }
      __syncthreads();
    }
    // ---- final stage on the tile: hard mask, then BCs
    // (rk3_simple_new.py:332-336, 264-279).  Boundary cells copy the masked
    // value of an interior source; each source writes its own targets so
    // every boundary cell is written exactly once.
#pragma unroll
    for (int it = 0; it < NT; ++it) {
      const int idx = tid + it * NTH;
      if (idx >= TX * TY) break;
      const int lx = H + idx / TY, ly = H + idx % TY;
      const int gx = gx0 + lx, gy = gy0 + ly;
      if (gx >= nx || gy >= ny) continue;
      const long g = (long)gx * ny + gy;
      if (gx == 0) {
        const bool wall = (gy == 0 || gy == ny - 1);
        a.u_out[g] = (wall && !a.slip) ? T(0) : a.U_inf;
        a.v_out[g] = T(0);
        if constexpr (DIV) a.rhs[g] = T(0);
        continue;
      }
      const bool interior = gx < nx - 1 && gy > 0 && gy < ny - 1;
      if (!interior) {
        if constexpr (DIV) a.rhs[g] = T(0);  // zero on every boundary cell
        continue;
      }
      const T m0 = rmask[it];
      T fu_, fv_;
      if constexpr (DIV) {
        constexpr int FW = TY + 2;
        const int f = (idx / TY + 1) * FW + (idx % TY + 1);
        fu_ = s_fu[f];
        fv_ = s_fv[f];
        // neighbours with the values the BCs give them: inlet u = U_inf,
        // outlet row = copy of row nx-2, wall v = 0
        const T uE = (gx == nx - 2) ? fu_ : s_fu[f + FW];
        const T uW = (gx == 1) ? a.U_inf : s_fu[f - FW];
        const T vN = (gy == ny - 2) ? T(0) : s_fv[f + 1];
        const T vS = (gy == 1) ? T(0) : s_fv[f - 1];
        const T d = rk_div(uE, uW, vN, vS, a.inv_2dx, a.inv_2dy);
        a.rhs[g] = a.rhs_masked ? d * m0 * a.inv_dt : d * a.inv_dt;
      } else {
        T nu_, nv_;
        stage(lx * RW + ly, gx, gy, nu_, nv_);
        fu_ = masked(nu_, m0);
        fv_ = masked(nv_, m0);
      }
      a.u_out[g] = fu_;
      a.v_out[g] = fv_;
      const T uw = a.slip ? fu_ : T(0);
      const bool out_row = (gx == nx - 2);
      if (out_row) {
        a.u_out[g + ny] = fu_;
        a.v_out[g + ny] = fv_;
      }
      if (gy == 1) {
        a.u_out[g - 1] = uw;
        a.v_out[g - 1] = T(0);
        if (out_row) { a.u_out[g + ny - 1] = uw; a.v_out[g + ny - 1] = T(0); }
      }
      if (gy == ny - 2) {
        a.u_out[g + 1] = uw;
        a.v_out[g + 1] = T(0);
        if (out_row) { a.u_out[g + ny + 1] = uw; a.v_out[g + ny + 1] = T(0); }
      }
    }
  }
}


// rhs at (i, j): divergence_nonperiodic(u*, v*) [* mask] / dt
template <typename T>
__device__ __forceinline__ void div_rhs_cell(const T* __restrict__ us, const T* __restrict__ vs,
                                             const T* __restrict__ mask, T* __restrict__ rhs,
                                             int nx, int ny, int i, int j, T inv_2dx, T inv_2dy,
                                             T inv_dt) {
  const long g = (long)i * ny + j;
  T r = T(0);
  if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
    const T d = rk_div(us[g + ny], us[g - ny], vs[g + 1], vs[g - 1], inv_2dx, inv_2dy);
    r = mask ? d * mask[g] * inv_dt : d * inv_dt;
  }
  rhs[g] = r;
}

// projected velocity of an interior cell from its stencil operands
// (step.py:257-264, 305-349, 364-377): the pressure gradient, with its
// normal component removed where the mask gradient marks an interface,
// subtracted from u*, v*, then masked.  One function for the per-cell and
// the per-pair kernels, with its multiply-adds spelled out, so that both give
// the same bits.
template <typename T>
__device__ __forceinline__ void project_values(T pxp, T pxm, T pyp, T pym, T mxp, T mxm, T myp,
                                               T mym, T usc, T vsc, T m0, T dt, T inv_2dx,
                                               T inv_2dy, T iface_thr, T* un, T* vn) {
#pragma clang fp contract(off)
  const T gx0 = (pxp - pxm) * inv_2dx;
  const T gy0 = (pyp - pym) * inv_2dy;
  const T mx = (mxp - mxm) * inv_2dx;
  const T my = (myp - mym) * inv_2dy;
  const T gm = sqrt(fma(mx, mx, my * my));
  T gx = gx0, gy = gy0;
  if (gm > iface_thr) {
    const T nxv = mx / (gm + T(1e-8));
    const T nyv = my / (gm + T(1e-8));
    const T dpn = fma(gx0, nxv, gy0 * nyv);
    gx = fma(-dpn, nxv, gx0);
    gy = fma(-dpn, nyv, gy0);
  }
  *un = masked(fma(-dt, gx, usc), m0);
  *vn = masked(fma(-dt, gy, vsc), m0);
}

// projection at (i, j) (step.py:257-264, 305-349, 364-377, 394-397); the
// boundary cells are written by their interior source
template <typename T>
__device__ __forceinline__ void project_cell(const T* __restrict__ us, const T* __restrict__ vs,
                                             const T* __restrict__ p, const T* __restrict__ mask,
                                             T* __restrict__ uo, T* __restrict__ vo, int nx,
                                             int ny, int i, int j, T dt, T inv_2dx, T inv_2dy,
                                             T iface_thr, T U_inf) {
  const long g = (long)i * ny + j;
  if (i == 0) {
    // inlet: u = U_inf (walls copy it), v = 0, then the final mask
    const T m0 = mask[g];
    uo[g] = masked(U_inf, m0);
    vo[g] = T(0);
    return;
  }
  if (!(i < nx - 1 && j > 0 && j < ny - 1)) return;  // written by a source
  T un, vn;
  project_values(p[g + ny], p[g - ny], p[g + 1], p[g - 1], mask[g + ny], mask[g - ny],
                 mask[g + 1], mask[g - 1], us[g], vs[g], mask[g], dt, inv_2dx, inv_2dy,
                 iface_thr, &un, &vn);
  uo[g] = un;  // second masking is idempotent on the source cell
  vo[g] = vn;
  const bool out_row = (i == nx - 2);
  if (out_row) {
    uo[g + ny] = masked(un, mask[g + ny]);
    vo[g + ny] = masked(vn, mask[g + ny]);
  }
  if (j == 1) {
    uo[g - 1] = masked(un, mask[g - 1]);
    vo[g - 1] = T(0);
    if (out_row) { uo[g + ny - 1] = masked(un, mask[g + ny - 1]); vo[g + ny - 1] = T(0); }
  }
  if (j == ny - 2) {
    uo[g + 1] = masked(un, mask[g + 1]);
    vo[g + 1] = T(0);
    if (out_row) { uo[g + ny + 1] = masked(un, mask[g + ny + 1]); vo[g + ny + 1] = T(0); }
  }
}

template <typename T> struct Pair;
template <> struct Pair<float> { using V = float2; };
template <> struct Pair<double> { using V = double2; };

// project_cell for the two cells (i, j) and (i, j + 1) of an even j on a grid
// of even ny >= 4, with the six arrays aligned to a pair: the centre rows
// move as one 2-vector per array, and the wall cells j = 0 and ny - 1 (each
// in the pair of its source) and the outlet row (written with the pairs of
// row nx - 2) come out as pairs too.  Same values as project_cell.
template <typename T>
__device__ __forceinline__ void project_pair(const T* __restrict__ us, const T* __restrict__ vs,
                                             const T* __restrict__ p, const T* __restrict__ mask,
                                             T* __restrict__ uo, T* __restrict__ vo, int nx,
                                             int ny, int i, int j, T dt, T inv_2dx, T inv_2dy,
                                             T iface_thr, T U_inf) {
  using V = typename Pair<T>::V;
  const long g = (long)i * ny + j;
  const V mc = *reinterpret_cast<const V*>(mask + g);
  if (i == 0) {
    V u2, v2;
    u2.x = masked(U_inf, mc.x);
    u2.y = masked(U_inf, mc.y);
    v2.x = v2.y = T(0);
    *reinterpret_cast<V*>(uo + g) = u2;
    *reinterpret_cast<V*>(vo + g) = v2;
    return;
  }
  if (i >= nx - 1) return;  // outlet row: written by row nx - 2
  // g - 1 and g + 2 stay inside the grid: rows i - 1 and i + 1 exist
  const V pxp = *reinterpret_cast<const V*>(p + g + ny);
  const V pxm = *reinterpret_cast<const V*>(p + g - ny);
  const V pc = *reinterpret_cast<const V*>(p + g);
  const V mxp = *reinterpret_cast<const V*>(mask + g + ny);
  const V mxm = *reinterpret_cast<const V*>(mask + g - ny);
  const V usc = *reinterpret_cast<const V*>(us + g);
  const V vsc = *reinterpret_cast<const V*>(vs + g);
  const T pym = p[g - 1], pyp = p[g + 2], mym = mask[g - 1], myp = mask[g + 2];
  T ua, va, ub, vb;
  project_values(pxp.x, pxm.x, pc.y, pym, mxp.x, mxm.x, mc.y, mym, usc.x, vsc.x, mc.x, dt,
                 inv_2dx, inv_2dy, iface_thr, &ua, &va);
  project_values(pxp.y, pxm.y, pyp, pc.x, mxp.y, mxm.y, myp, mc.x, usc.y, vsc.y, mc.y, dt,
                 inv_2dx, inv_2dy, iface_thr, &ub, &vb);
  // the wall cell of an edge pair takes its neighbour's u (v = 0)
  const bool lo = (j == 0), hi = (j == ny - 2);
  const T su = lo ? ub : ua, sv = lo ? T(0) : va;  // source of cell j
  const T tu = hi ? ua : ub, tv = hi ? T(0) : vb;  // source of cell j + 1
  V u2, v2;
  u2.x = lo ? masked(su, mc.x) : su;
  u2.y = hi ? masked(tu, mc.y) : tu;
  v2.x = sv;
  v2.y = tv;
  *reinterpret_cast<V*>(uo + g) = u2;
  *reinterpret_cast<V*>(vo + g) = v2;
  if (i == nx - 2) {
    u2.x = masked(su, mxp.x);
    u2.y = masked(tu, mxp.y);
    v2.x = masked(sv, mxp.x);
    v2.y = masked(tv, mxp.y);
    *reinterpret_cast<V*>(uo + g + ny) = u2;
    *reinterpret_cast<V*>(vo + g + ny) = v2;
  }
}

}  // namespace aj
