// Step kernels: fused RK predictor, Smagorinsky eddy viscosity, pressure
// right-hand side and projection (collocated grid, von Karman BCs).
//
// All kernels are HBM-bound stencils.  The RK predictor fuses its 2 or 3
// stages into one pass: a workgroup loads u, v on its (TX x TY) tile plus a
// halo of (stages x per-stage radius) cells into LDS, evaluates each stage on
// a region that shrinks by the stencil radius, and writes only the tile, so
// HBM sees u, v, chi-source, mask read once and u*, v* written once.
// Neighbour reads clamp the *global* index at the domain edge, which is the
// reference's jnp.pad(mode='edge') applied to every stage field.
#include "aj_step_kernels.h"

#include <algorithm>
#include <cstdio>

#include "aj_rk_rc.h"

namespace aj {

template <typename T, int TX, int TY, bool FAST, bool HYPER, int NTH, bool DIV>
__global__ __launch_bounds__(NTH) void k_rk_fused(RkArgs<T> a) {
  extern __shared__ unsigned char smem_raw[];
  int bx, by;
  tile_block(&bx, &by);
  rk_tile<T, TX, TY, FAST, HYPER, NTH, DIV>(a, by * TX, bx * TY,
                                            reinterpret_cast<T*>(smem_raw));
}

template <typename T, int TX, int TY, bool FAST, bool HYPER, bool DIV = false>
static int launch_rk(const RkArgs<T>& a, hipStream_t st) {
  constexpr int NTH = 512;
  constexpr size_t N = RkTileShape<TX, TY, FAST, HYPER, DIV>::N;
  const size_t smem = N * sizeof(T);
  static_assert(N * sizeof(T) <= 160 * 1024, "LDS budget");
  if (smem > 65536) {
    // dynamic LDS beyond 64 KiB has to be enabled per kernel (gfx950: 160 KiB)
    static bool raised = false;
    if (!raised) {
      AJ_HIP(hipFuncSetAttribute((const void*)k_rk_fused<T, TX, TY, FAST, HYPER, NTH, DIV>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem));
      raised = true;
    }
  }
  dim3 grid(cdiv(a.ny, TY), cdiv(a.nx, TX));
  hipLaunchKernelGGL((k_rk_fused<T, TX, TY, FAST, HYPER, NTH, DIV>), grid, dim3(NTH), smem, st,
                     a);
  AJ_LAUNCH_CHECK("k_rk_fused");
  return AJ_OK;
}

// register-column predictor (FP32, no hyperviscosity): W waves x R rows per
// tile; refused (a failing tuner candidate) for odd or tiny grids
template <int W, int R, bool FAST, bool DIV>
static int launch_rk_rc(const RkArgs<float>& a, hipStream_t st) {
  if ((a.ny & 1) || a.ny < 4 || a.nx < 4)
    return fail(AJ_ERR_INVALID, "register-column RK predictor: odd or tiny grid");
  RkRcArgs r;
  r.u = a.u; r.v = a.v; r.mask = a.mask; r.sdf = a.sdf;
  r.nu_sgs = a.nu_sgs; r.nu_field = a.nu_field;
  r.u_out = a.u_out; r.v_out = a.v_out; r.rhs = a.rhs;
  r.nx = a.nx; r.ny = a.ny;
  r.dt = a.dt; r.nu = a.nu; r.inv_dx = a.inv_dx; r.inv_dy = a.inv_dy;
  r.inv_dx2 = a.inv_dx2; r.inv_dy2 = a.inv_dy2; r.U_inf = a.U_inf; r.inv_eta = a.inv_eta;
  r.sdf_scale = a.sdf_scale; r.inv_2dx = a.inv_2dx; r.inv_2dy = a.inv_2dy; r.inv_dt = a.inv_dt;
  r.slip = a.slip; r.rhs_masked = a.rhs_masked;
  using G = RkRcGeo<W, R, FAST, DIV>;
  const int XT = G::x_tiles(a.nx), YT = a.ny <= 128 ? 1 : (int)cdiv(a.ny, G::UY);
  hipLaunchKernelGGL((k_rk_rc<W, R, FAST, DIV>), dim3(8 * cdiv((long)XT * YT, 8)), dim3(W * 64),
                     0, st, r, XT, YT);
  AJ_LAUNCH_CHECK("k_rk_rc");
  return AJ_OK;
}

template <typename T>
static int rk_step_t(const void* u, const void* v, const void* mask, const void* sdf,
                     const void* nu_field, const void* nu_sgs, void* u_out, void* v_out, int nx,
                     int ny, const aj_rk_params* p, hipStream_t st, void* rhs = nullptr,
                     int rhs_masked = 1) {
  RkArgs<T> a;
  a.u = (const T*)u; a.v = (const T*)v; a.mask = (const T*)mask;
  a.sdf = (const T*)sdf; a.nu_sgs = (const T*)nu_sgs; a.nu_field = (const T*)nu_field;
  a.u_out = (T*)u_out; a.v_out = (T*)v_out;
  a.rhs = (T*)rhs;
  a.rhs_masked = rhs_masked;
  a.inv_2dx = T(1.0 / (2.0 * p->dx));
  a.inv_2dy = T(1.0 / (2.0 * p->dy));
  a.inv_dt = T(1.0 / p->dt);
  a.nx = nx; a.ny = ny;
  a.dt = T(p->dt); a.nu = T(p->nu);
  a.inv_dx = T(1.0 / p->dx); a.inv_dy = T(1.0 / p->dy);
  a.inv_dx2 = T(1.0 / (p->dx * p->dx)); a.inv_dy2 = T(1.0 / (p->dy * p->dy));
  a.U_inf = T(p->U_inf); a.inv_eta = T(1.0 / p->brinkman_eta);
  a.sdf_scale = T(1.0 / (0.1 * p->dx));
  a.hyper = T(p->nu_hyper_ratio * (p->dx * p->dy));
  a.slip = p->slip_walls ? 1 : 0;
  const bool hyper = p->nu_hyper_ratio > 0.0;
  const bool fast = p->fast_mode != 0;
  if (rhs) {
    // RK predictor with the pressure rhs in its epilogue (no hyperviscosity:
    // the GUI step that uses it keeps the separate rhs launch)
    auto launch = [&](int c, hipStream_t s) -> int {
      if constexpr (sizeof(T) == 8) {
        if (c == 1) return fast ? launch_rk<T, 8, 64, true, false, true>(a, s)
                                : launch_rk<T, 8, 64, false, false, true>(a, s);
        return fast ? launch_rk<T, 4, 64, true, false, true>(a, s)
                    : launch_rk<T, 4, 64, false, false, true>(a, s);
      } else {
        switch (c) {
          case 1: return fast ? launch_rk<T, 8, 64, true, false, true>(a, s)
                              : launch_rk<T, 8, 64, false, false, true>(a, s);
          case 2: return fast ? launch_rk<T, 32, 64, true, false, true>(a, s)
                              : launch_rk<T, 32, 64, false, false, true>(a, s);
          case 3: return fast ? launch_rk_rc<8, 4, true, true>(a, s)
                              : launch_rk_rc<8, 4, false, true>(a, s);
          case 4: return fast ? launch_rk_rc<8, 2, true, true>(a, s)
                              : launch_rk_rc<8, 2, false, true>(a, s);
          default: return fast ? launch_rk<T, 16, 64, true, false, true>(a, s)
                               : launch_rk<T, 16, 64, false, false, true>(a, s);
        }
      }
    };
    char key[96];
    std::snprintf(key, sizeof(key), "rkdiv/%s/%s/%dx%d", sizeof(T) == 8 ? "f64" : "f32",
                  fast ? "rk2" : "rk3", nx, ny);
    // u, v, chi source, mask (+ nu_sgs) read; u*, v*, rhs written
    ProfScope prof(AJ_PROF_RK, (double)nx * ny * sizeof(T) * (7 + (a.nu_sgs ? 1 : 0)), st);
    int pick = 0;
    return tune_pick(key, sizeof(T) == 8 ? 2 : 5, launch, st, &pick);
  }
  // tile candidates (rows x columns), tuned per grid shape and variant
  auto launch = [&](int c, hipStream_t s) -> int {
    if constexpr (sizeof(T) == 8) {
      if (hyper) return fast ? launch_rk<T, 4, 32, true, true>(a, s)
                             : launch_rk<T, 4, 32, false, true>(a, s);
      switch (c) {
        case 1: return fast ? launch_rk<T, 4, 64, true, false>(a, s)
                            : launch_rk<T, 4, 64, false, false>(a, s);
        case 2: return fast ? launch_rk<T, 16, 64, true, false>(a, s)
                            : launch_rk<T, 16, 64, false, false>(a, s);
        default: return fast ? launch_rk<T, 8, 64, true, false>(a, s)
                             : launch_rk<T, 8, 64, false, false>(a, s);
      }
    } else {
      if (hyper) return fast ? launch_rk<T, 8, 64, true, true>(a, s)
                             : launch_rk<T, 8, 64, false, true>(a, s);
      switch (c) {
        case 1: return fast ? launch_rk<T, 8, 64, true, false>(a, s)
                            : launch_rk<T, 8, 64, false, false>(a, s);
        case 2: return fast ? launch_rk<T, 32, 64, true, false>(a, s)
                            : launch_rk<T, 32, 64, false, false>(a, s);
        case 3: return fast ? launch_rk<T, 8, 128, true, false>(a, s)
                            : launch_rk<T, 8, 128, false, false>(a, s);
        case 4: return fast ? launch_rk_rc<8, 4, true, false>(a, s)
                            : launch_rk_rc<8, 4, false, false>(a, s);
        case 5: return fast ? launch_rk_rc<8, 2, true, false>(a, s)
                            : launch_rk_rc<8, 2, false, false>(a, s);
        default: return fast ? launch_rk<T, 16, 64, true, false>(a, s)
                             : launch_rk<T, 16, 64, false, false>(a, s);
      }
    }
  };
  const int ncand = hyper ? 1 : (sizeof(T) == 8 ? 3 : 6);
  char key[96];
  std::snprintf(key, sizeof(key), "rk/%s/%s%s/%dx%d", sizeof(T) == 8 ? "f64" : "f32",
                fast ? "rk2" : "rk3", hyper ? "+hyper" : "", nx, ny);
  // algorithmic bytes: u, v, chi source, mask (+ nu_sgs) read, u*, v* written
  ProfScope prof(AJ_PROF_RK, (double)nx * ny * sizeof(T) * (6 + (a.nu_sgs ? 1 : 0)), st);
  int pick = 0;
  return tune_pick(key, ncand, launch, st, &pick);
}

// --------------------------------------------------------------------------
// Smagorinsky (solver/les_models.py:22-54, 179-200)
// --------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_les(const T* __restrict__ u, const T* __restrict__ v,
                                             T* __restrict__ nsgs, T* __restrict__ ntot, T nu,
                                             int nx, int ny, T inv_2dx, T inv_2dy, T coef) {
  int i, j;
  row_block(&i, &j);
  j += threadIdx.x;
  if (j >= ny) return;
  const long c = (long)i * ny;
  const T s = les_cell(u, v, nx, ny, i, j, inv_2dx, inv_2dy, coef);
  nsgs[c + j] = s;
  if (ntot) ntot[c + j] = nu + s;  // nu_total = nu + nu_sgs (step_handlers.py:76)
}

// --------------------------------------------------------------------------
// rhs = divergence_nonperiodic(u*, v*) * mask / dt (operators.py:81-96,
// step.py:194-199): zero on every boundary cell.
// --------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_div_rhs(const T* __restrict__ us, const T* __restrict__ vs,
                                                 const T* __restrict__ mask, T* __restrict__ rhs,
                                                 int nx, int ny, T inv_2dx, T inv_2dy, T inv_dt) {
  int i, j;
  row_block(&i, &j);
  j += threadIdx.x;
  if (j >= ny) return;
  div_rhs_cell(us, vs, mask, rhs, nx, ny, i, j, inv_2dx, inv_2dy, inv_dt);
}

// --------------------------------------------------------------------------
// Projection (step.py:257-264, 305-349, 364-377, 394-397)
// --------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(256) void k_project(const T* __restrict__ us, const T* __restrict__ vs,
                                                 const T* __restrict__ p, const T* __restrict__ mask,
                                                 T* __restrict__ uo, T* __restrict__ vo, int nx,
                                                 int ny, T dt, T inv_2dx, T inv_2dy, T iface_thr,
                                                 T U_inf) {
  int i, j;
  row_block(&i, &j);
  j += threadIdx.x;
  if (j >= ny) return;
  project_cell(us, vs, p, mask, uo, vo, nx, ny, i, j, dt, inv_2dx, inv_2dy, iface_thr, U_inf);
}

// k_project over pairs of cells (grid row_grid(nx, ny / 2), ny even)
template <typename T>
__global__ __launch_bounds__(256) void k_project_pair(
    const T* __restrict__ us, const T* __restrict__ vs, const T* __restrict__ p,
    const T* __restrict__ mask, T* __restrict__ uo, T* __restrict__ vo, int nx, int ny, T dt,
    T inv_2dx, T inv_2dy, T iface_thr, T U_inf) {
  int i, q;
  row_block(&i, &q);
  q += threadIdx.x;
  if (2 * q >= ny) return;
  project_pair(us, vs, p, mask, uo, vo, nx, ny, i, 2 * q, dt, inv_2dx, inv_2dy, iface_thr, U_inf);
}

// one block per 256-column segment of a row; blocks are renumbered per XCD
// (row_block) so that the +-row stencil reads of a band of rows share an L2
static dim3 row_grid(int nx, int ny) { return dim3(cdiv(ny, 256), nx); }

// pairs need an even row length and pair-aligned arrays
static bool pair_ok(int ny, size_t elem, const void* const* ptrs, int n) {
  if (ny & 1) return false;
  for (int k = 0; k < n; ++k)
    if ((uintptr_t)ptrs[k] % (2 * elem)) return false;
  return true;
}

static int check_grid(int nx, int ny) {
  if (nx < 4 || ny < 4) return fail(AJ_ERR_INVALID, "grid must be at least 4x4");
  if ((long)nx * ny > (1L << 31)) return fail(AJ_ERR_INVALID, "grid too large");
  return AJ_OK;
}

// Projection tail of the GUI step (_step_collocated,
// solver/simulation/step_handlers.py:131-176): grad_*_nonperiodic of p,
// dp/dx[0, :] = 0 (the corner taper of solver/pure_functions/
// corner_smoothing.py:39-61 then scales zeros), u = u* - dt grad p; inlet
// u[0, 1:-1] = U_in, v[0, :] = 0; outlet row nx-1 copies row nx-2; no-slip
// walls u = v = 0 on j = 0, ny-1 (corners included); then the hard mask.
// Every cell computes its own value (row nx-1 recomputes row nx-2's
// correction, whose gradients are central).
template <typename T>
__global__ __launch_bounds__(256) void k_project_gui(const T* __restrict__ us,
                                                     const T* __restrict__ vs,
                                                     const T* __restrict__ p,
                                                     const T* __restrict__ mask,
                                                     T* __restrict__ uo, T* __restrict__ vo,
                                                     int nx, int ny, T dt, T inv_2dx, T inv_2dy,
                                                     T U_in) {
  int i, j;
  row_block(&i, &j);
  j += threadIdx.x;
  if (j >= ny) return;
  const long g = (long)i * ny + j;
  T un = T(0), vn = T(0);
  if (j > 0 && j < ny - 1) {
    if (i == 0) {
      un = U_in;
    } else {
      const long s = (long)(i == nx - 1 ? nx - 2 : i) * ny + j;
      const T gx = (p[s + ny] - p[s - ny]) * inv_2dx;
      const T gy = (p[s + 1] - p[s - 1]) * inv_2dy;
      un = us[s] - dt * gx;
      vn = vs[s] - dt * gy;
    }
  }
  const T m0 = mask[g];
  uo[g] = masked(un, m0);
  vo[g] = masked(vn, m0);
}

}  // namespace aj

using namespace aj;

static int rk_entry(int dtype, const void* u, const void* v, const void* mask, const void* sdf,
                    const void* nu_field, const void* nu_sgs, void* u_out, void* v_out, int nx,
                    int ny, const aj_rk_params* prm, void* stream) {
  if (!u || !v || !mask || !u_out || !v_out || !prm) return fail(AJ_ERR_INVALID, "aj_rk_step: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  if (!(prm->brinkman_eta > 0)) return fail(AJ_ERR_INVALID, "aj_rk_step: brinkman_eta must be > 0");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == AJ_F32)
    return rk_step_t<float>(u, v, mask, sdf, nu_field, nu_sgs, u_out, v_out, nx, ny, prm, st);
  if (dtype == AJ_F64)
    return rk_step_t<double>(u, v, mask, sdf, nu_field, nu_sgs, u_out, v_out, nx, ny, prm, st);
  return fail(AJ_ERR_INVALID, "aj_rk_step: bad dtype");
}

extern "C" int aj_rk_step(int dtype, const void* u, const void* v, const void* mask,
                          const void* sdf, const void* nu_sgs, void* u_out, void* v_out,
                          int nx, int ny, const aj_rk_params* prm, void* stream) {
  return rk_entry(dtype, u, v, mask, sdf, nullptr, nu_sgs, u_out, v_out, nx, ny, prm, stream);
}

extern "C" int aj_rk_step_rhs(int dtype, const void* u, const void* v, const void* mask,
                              const void* sdf, const void* nu_sgs, void* u_out, void* v_out,
                              void* rhs, int rhs_masked, int nx, int ny, const aj_rk_params* prm,
                              void* stream) {
  if (!u || !v || !mask || !u_out || !v_out || !rhs || !prm)
    return fail(AJ_ERR_INVALID, "aj_rk_step_rhs: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  if (!(prm->brinkman_eta > 0)) return fail(AJ_ERR_INVALID, "aj_rk_step_rhs: brinkman_eta must be > 0");
  if (prm->nu_hyper_ratio > 0.0)
    return fail(AJ_ERR_UNSUPPORTED, "aj_rk_step_rhs: hyperviscosity needs aj_rk_step + aj_divergence_rhs");
  if (rhs == u_out || rhs == v_out || rhs == u || rhs == v || rhs == mask)
    return fail(AJ_ERR_INVALID, "aj_rk_step_rhs: rhs must not alias another field");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == AJ_F32)
    return rk_step_t<float>(u, v, mask, sdf, nullptr, nu_sgs, u_out, v_out, nx, ny, prm, st, rhs,
                            rhs_masked ? 1 : 0);
  if (dtype == AJ_F64)
    return rk_step_t<double>(u, v, mask, sdf, nullptr, nu_sgs, u_out, v_out, nx, ny, prm, st, rhs,
                             rhs_masked ? 1 : 0);
  return fail(AJ_ERR_INVALID, "aj_rk_step_rhs: bad dtype");
}

extern "C" int aj_rk_step_nu_field(int dtype, const void* u, const void* v, const void* mask,
                                   const void* sdf, const void* nu_field, const void* nu_sgs,
                                   void* u_out, void* v_out, int nx, int ny,
                                   const aj_rk_params* prm, void* stream) {
  if (!nu_field) return fail(AJ_ERR_INVALID, "aj_rk_step_nu_field: null nu_field");
  return rk_entry(dtype, u, v, mask, sdf, nu_field, nu_sgs, u_out, v_out, nx, ny, prm, stream);
}

static int les_entry(int dtype, const void* u, const void* v, double nu, void* nu_sgs,
                     void* nu_total, int nx, int ny, double dx, double dy, double delta,
                     double c_s, void* stream) {
  if (!u || !v || !nu_sgs) return fail(AJ_ERR_INVALID, "aj_les_smagorinsky: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const double coef = (c_s * delta) * (c_s * delta);
  if (dtype == AJ_F32) {
    hipLaunchKernelGGL(k_les<float>, row_grid(nx, ny), dim3(256), 0, st, (const float*)u,
                       (const float*)v, (float*)nu_sgs, (float*)nu_total, float(nu), nx, ny,
                       float(1.0 / (2.0 * dx)), float(1.0 / (2.0 * dy)), float(coef));
  } else if (dtype == AJ_F64) {
    hipLaunchKernelGGL(k_les<double>, row_grid(nx, ny), dim3(256), 0, st, (const double*)u,
                       (const double*)v, (double*)nu_sgs, (double*)nu_total, nu, nx, ny,
                       1.0 / (2.0 * dx), 1.0 / (2.0 * dy), coef);
  } else {
    return fail(AJ_ERR_INVALID, "aj_les_smagorinsky: bad dtype");
  }
  AJ_LAUNCH_CHECK("k_les");
  return AJ_OK;
}

extern "C" int aj_les_smagorinsky(int dtype, const void* u, const void* v, void* nu_sgs, int nx,
                                  int ny, double dx, double dy, double delta, double c_s,
                                  void* stream) {
  return les_entry(dtype, u, v, 0.0, nu_sgs, nullptr, nx, ny, dx, dy, delta, c_s, stream);
}

extern "C" int aj_les_smagorinsky_total(int dtype, const void* u, const void* v, double nu,
                                        void* nu_sgs, void* nu_total, int nx, int ny, double dx,
                                        double dy, double delta, double c_s, void* stream) {
  if (!nu_total) return fail(AJ_ERR_INVALID, "aj_les_smagorinsky_total: null nu_total");
  return les_entry(dtype, u, v, nu, nu_sgs, nu_total, nx, ny, dx, dy, delta, c_s, stream);
}

extern "C" int aj_divergence_rhs(int dtype, const void* us, const void* vs, const void* mask,
                                 void* rhs, int nx, int ny, double dx, double dy, double dt,
                                 void* stream) {
  if (!us || !vs || !rhs) return fail(AJ_ERR_INVALID, "aj_divergence_rhs: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  hipStream_t st = (hipStream_t)stream;
  // u*, v* (+ mask) read, rhs written
  ProfScope prof(AJ_PROF_RHS, (double)nx * ny * (dtype == AJ_F64 ? 8 : 4) * (mask ? 4 : 3), st);
  if (dtype == AJ_F32) {
    hipLaunchKernelGGL(k_div_rhs<float>, row_grid(nx, ny), dim3(256), 0, st, (const float*)us,
                       (const float*)vs, (const float*)mask, (float*)rhs, nx, ny,
                       float(1.0 / (2.0 * dx)), float(1.0 / (2.0 * dy)), float(1.0 / dt));
  } else if (dtype == AJ_F64) {
    hipLaunchKernelGGL(k_div_rhs<double>, row_grid(nx, ny), dim3(256), 0, st, (const double*)us,
                       (const double*)vs, (const double*)mask, (double*)rhs, nx, ny,
                       1.0 / (2.0 * dx), 1.0 / (2.0 * dy), 1.0 / dt);
  } else {
    return fail(AJ_ERR_INVALID, "aj_divergence_rhs: bad dtype");
  }
  AJ_LAUNCH_CHECK("k_div_rhs");
  return AJ_OK;
}

// periodic divergence grad_x(u) + grad_y(v) with jnp.roll wrap
// (solver/operators.py:11-14, 35-38, 76-78)
template <typename T>
__global__ __launch_bounds__(256) void k_div_periodic(const T* __restrict__ u,
                                                      const T* __restrict__ v, T* __restrict__ div,
                                                      int nx, int ny, T inv_2dx, T inv_2dy) {
  int i, j;
  row_block(&i, &j);
  j += threadIdx.x;
  if (j >= ny) return;
  const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
  const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
  const long c = (long)i * ny;
  div[c + j] = (u[(long)ip * ny + j] - u[(long)im * ny + j]) * inv_2dx +
               (v[c + jp] - v[c + jm]) * inv_2dy;
}

extern "C" int aj_divergence_periodic(int dtype, const void* u, const void* v, void* div, int nx,
                                      int ny, double dx, double dy, void* stream) {
  if (!u || !v || !div) return fail(AJ_ERR_INVALID, "aj_divergence_periodic: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (dtype == AJ_F32) {
    hipLaunchKernelGGL(k_div_periodic<float>, row_grid(nx, ny), dim3(256), 0, st, (const float*)u,
                       (const float*)v, (float*)div, nx, ny, float(1.0 / (2.0 * dx)),
                       float(1.0 / (2.0 * dy)));
  } else if (dtype == AJ_F64) {
    hipLaunchKernelGGL(k_div_periodic<double>, row_grid(nx, ny), dim3(256), 0, st,
                       (const double*)u, (const double*)v, (double*)div, nx, ny,
                       1.0 / (2.0 * dx), 1.0 / (2.0 * dy));
  } else {
    return fail(AJ_ERR_INVALID, "aj_divergence_periodic: bad dtype");
  }
  AJ_LAUNCH_CHECK("k_div_periodic");
  return AJ_OK;
}

extern "C" int aj_project(int dtype, const void* us, const void* vs, const void* p,
                          const void* mask, void* uo, void* vo, int nx, int ny, double dt,
                          double dx, double dy, double U_inf, void* stream) {
  if (!us || !vs || !p || !mask || !uo || !vo) return fail(AJ_ERR_INVALID, "aj_project: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  hipStream_t st = (hipStream_t)stream;
  // u*, v*, p, mask read; u, v written
  ProfScope prof(AJ_PROF_PROJECT, (double)nx * ny * (dtype == AJ_F64 ? 8 : 4) * 6, st);
  const void* ptrs[6] = {us, vs, p, mask, uo, vo};
  auto launch = [&](auto zero) {
    using T = decltype(zero);
    const T* a[4] = {(const T*)us, (const T*)vs, (const T*)p, (const T*)mask};
    const T c[5] = {T(dt), T(1.0 / (2.0 * dx)), T(1.0 / (2.0 * dy)), T(0.1 / dx), T(U_inf)};
    if (pair_ok(ny, sizeof(T), ptrs, 6))
      hipLaunchKernelGGL(k_project_pair<T>, row_grid(nx, ny / 2), dim3(256), 0, st, a[0], a[1],
                         a[2], a[3], (T*)uo, (T*)vo, nx, ny, c[0], c[1], c[2], c[3], c[4]);
    else
      hipLaunchKernelGGL(k_project<T>, row_grid(nx, ny), dim3(256), 0, st, a[0], a[1], a[2],
                         a[3], (T*)uo, (T*)vo, nx, ny, c[0], c[1], c[2], c[3], c[4]);
  };
  if (dtype == AJ_F32) {
    launch(0.f);
  } else if (dtype == AJ_F64) {
    launch(0.0);
  } else {
    return fail(AJ_ERR_INVALID, "aj_project: bad dtype");
  }
  AJ_LAUNCH_CHECK("k_project");
  return AJ_OK;
}

extern "C" int aj_project_collocated(int dtype, const void* us, const void* vs, const void* p,
                                     const void* mask, void* uo, void* vo, int nx, int ny,
                                     double dt, double dx, double dy, double U_in, void* stream) {
  if (!us || !vs || !p || !mask || !uo || !vo)
    return fail(AJ_ERR_INVALID, "aj_project_collocated: null pointer");
  if (int rc = check_grid(nx, ny)) return rc;
  hipStream_t st = (hipStream_t)stream;
  ProfScope prof(AJ_PROF_PROJECT, (double)nx * ny * (dtype == AJ_F64 ? 8 : 4) * 6, st);
  if (dtype == AJ_F32) {
    hipLaunchKernelGGL(k_project_gui<float>, row_grid(nx, ny), dim3(256), 0, st,
                       (const float*)us, (const float*)vs, (const float*)p, (const float*)mask,
                       (float*)uo, (float*)vo, nx, ny, float(dt), float(1.0 / (2.0 * dx)),
                       float(1.0 / (2.0 * dy)), float(U_in));
  } else if (dtype == AJ_F64) {
    hipLaunchKernelGGL(k_project_gui<double>, row_grid(nx, ny), dim3(256), 0, st,
                       (const double*)us, (const double*)vs, (const double*)p,
                       (const double*)mask, (double*)uo, (double*)vo, nx, ny, dt,
                       1.0 / (2.0 * dx), 1.0 / (2.0 * dy), U_in);
  } else {
    return fail(AJ_ERR_INVALID, "aj_project_collocated: bad dtype");
  }
  AJ_LAUNCH_CHECK("k_project_gui");
  return AJ_OK;
}
