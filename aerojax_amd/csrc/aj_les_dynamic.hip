// Dynamic Smagorinsky eddy viscosity (solver/les_models.py:57-176),
// collocated grid, in four passes of one cell per thread:
//   strain : |S| and |S| S_ij                           (periodic differences)
//   filter : the 5x5 periodic box filter of u, v, uu, uv, vv, |S|S_ij, summed
//            in the reference's scan order (x offset outer, y inner)
//   lm     : test-level strain of the filtered velocity, Leonard stress
//            L_ij = u~_i u~_j - (u_i u_j)~, M_ij = alpha^2 |S~| S~_ij -
//            (|S| S_ij)~, and the contractions L:M, M:M
//   final  : 3x3 interior averages of L:M and M:M (zero on the border),
//            C_d = clip(LM / (MM + 1e-8), 0, 0.2), nu_sgs = C_d delta^2 |S|
// The intermediate fields live in a stream-ordered scratch allocation.
#include "aj_common.h"

namespace aj {

template <typename T>
struct DynF {
  T *S, *Sxx, *Sxy, *Syy;                        // |S| and |S| S_ij
  T *ut, *vt, *uut, *uvt, *vvt, *SSxx, *SSxy, *SSyy;  // filtered
  T *lm, *mm;
};

template <typename T>
__device__ __forceinline__ void strain_at(const T* __restrict__ u, const T* __restrict__ v, int nx,
                                          int ny, int i, int j, T inv_2dx, T inv_2dy, T& S, T& xx,
                                          T& yy, T& xy) {
  const int ip = (i + 1 == nx) ? 0 : i + 1, im = (i == 0) ? nx - 1 : i - 1;
  const int jp = (j + 1 == ny) ? 0 : j + 1, jm = (j == 0) ? ny - 1 : j - 1;
  const long c = (long)i * ny;
  const T du_dx = (u[(long)ip * ny + j] - u[(long)im * ny + j]) * inv_2dx;
  const T du_dy = (u[c + jp] - u[c + jm]) * inv_2dy;
  const T dv_dx = (v[(long)ip * ny + j] - v[(long)im * ny + j]) * inv_2dx;
  const T dv_dy = (v[c + jp] - v[c + jm]) * inv_2dy;
  xx = du_dx;
  yy = dv_dy;
  xy = T(0.5) * (du_dy + dv_dx);
  S = sqrt(T(2) * (xx * xx + yy * yy + T(2) * xy * xy) + T(1e-10));
}

template <typename T>
__global__ __launch_bounds__(256) void k_dyn_strain(const T* __restrict__ u, const T* __restrict__ v,
                                                    int nx, int ny, T inv_2dx, T inv_2dy, DynF<T> f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ny) return;
  T S, xx, yy, xy;
  strain_at(u, v, nx, ny, i, j, inv_2dx, inv_2dy, S, xx, yy, xy);
  const long c = (long)i * ny + j;
  f.S[c] = S;
  f.Sxx[c] = S * xx;
  f.Sxy[c] = S * xy;
  f.Syy[c] = S * yy;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dyn_filter(const T* __restrict__ u, const T* __restrict__ v,
                                                    int nx, int ny, DynF<T> f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ny) return;
  T a[8] = {T(0), T(0), T(0), T(0), T(0), T(0), T(0), T(0)};
  for (int di = -2; di <= 2; ++di) {
    int x = i - di;
    x = (x < 0) ? x + nx : (x >= nx ? x - nx : x);
    for (int dj = -2; dj <= 2; ++dj) {
      int y = j - dj;
      y = (y < 0) ? y + ny : (y >= ny ? y - ny : y);
      const long q = (long)x * ny + y;
      const T uq = u[q], vq = v[q];
      a[0] += uq;
      a[1] += vq;
      a[2] += uq * uq;
      a[3] += uq * vq;
      a[4] += vq * vq;
      a[5] += f.Sxx[q];
      a[6] += f.Sxy[q];
      a[7] += f.Syy[q];
    }
  }
  const long c = (long)i * ny + j;
  const T k = T(25);
  f.ut[c] = a[0] / k;
  f.vt[c] = a[1] / k;
  f.uut[c] = a[2] / k;
  f.uvt[c] = a[3] / k;
  f.vvt[c] = a[4] / k;
  f.SSxx[c] = a[5] / k;
  f.SSxy[c] = a[6] / k;
  f.SSyy[c] = a[7] / k;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dyn_lm(int nx, int ny, T inv_2dx, T inv_2dy, T a2,
                                                DynF<T> f) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ny) return;
  T St, xx, yy, xy;
  strain_at<T>(f.ut, f.vt, nx, ny, i, j, inv_2dx, inv_2dy, St, xx, yy, xy);
  const long c = (long)i * ny + j;
  const T ut = f.ut[c], vt = f.vt[c];
  const T L11 = ut * ut - f.uut[c];
  const T L12 = ut * vt - f.uvt[c];
  const T L22 = vt * vt - f.vvt[c];
  const T M11 = a2 * St * xx - f.SSxx[c];
  const T M12 = a2 * St * xy - f.SSxy[c];
  const T M22 = a2 * St * yy - f.SSyy[c];
  f.lm[c] = L11 * M11 + L12 * M12 + L22 * M22;
  f.mm[c] = M11 * M11 + M12 * M12 + M22 * M22;
}

template <typename T>
__device__ __forceinline__ T avg9(const T* __restrict__ g, long c, int ny) {
  // f[1:-1,1:-1] + f[2:,1:-1] + f[:-2,1:-1] + f[1:-1,2:] + f[1:-1,:-2]
  // + f[2:,2:] + f[2:,:-2] + f[:-2,2:] + f[:-2,:-2], then / 9
  const long E = c + ny, W = c - ny;
  return (g[c] + g[E] + g[W] + g[c + 1] + g[c - 1] + g[E + 1] + g[E - 1] + g[W + 1] + g[W - 1]) /
         T(9);
}

template <typename T>
__global__ __launch_bounds__(256) void k_dyn_final(int nx, int ny, T dsq, T nu, DynF<T> f,
                                                   T* __restrict__ nsgs, T* __restrict__ ntot,
                                                   T* __restrict__ cd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x, i = blockIdx.y;
  if (j >= ny) return;
  const long c = (long)i * ny + j;
  T C = T(0);
  if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1) {
    const T LM = avg9(f.lm, c, ny), MM = avg9(f.mm, c, ny);
    C = LM / (MM + T(1e-8));
    C = C < T(0) ? T(0) : (C > T(0.2) ? T(0.2) : C);
  }
  const T s = C * dsq * f.S[c];
  nsgs[c] = s;
  if (ntot) ntot[c] = nu + s;  // nu_total = nu + nu_sgs (step_handlers.py:76)
  if (cd) cd[c] = C;
}

template <typename T>
static int dyn_launch(const T* u, const T* v, T nu, T* nsgs, T* ntot, T* cd, int nx, int ny,
                      double dx, double dy, double delta, double alpha, hipStream_t st) {
  const long n = (long)nx * ny;
  void* mem = nullptr;
  AJ_HIP(hipMallocAsync(&mem, (size_t)n * 14 * sizeof(T), st));
  T* w = (T*)mem;
  DynF<T> f;
  T** slots[14] = {&f.S, &f.Sxx, &f.Sxy, &f.Syy, &f.ut, &f.vt, &f.uut, &f.uvt, &f.vvt,
                   &f.SSxx, &f.SSxy, &f.SSyy, &f.lm, &f.mm};
  for (int k = 0; k < 14; ++k) *slots[k] = w + (long)k * n;
  const dim3 grid(cdiv(ny, 256), nx);
  const T i2x = T(1.0 / (2.0 * dx)), i2y = T(1.0 / (2.0 * dy));
  hipLaunchKernelGGL(k_dyn_strain<T>, grid, dim3(256), 0, st, u, v, nx, ny, i2x, i2y, f);
  AJ_LAUNCH_CHECK("k_dyn_strain");
  hipLaunchKernelGGL(k_dyn_filter<T>, grid, dim3(256), 0, st, u, v, nx, ny, f);
  AJ_LAUNCH_CHECK("k_dyn_filter");
  hipLaunchKernelGGL(k_dyn_lm<T>, grid, dim3(256), 0, st, nx, ny, i2x, i2y, T(alpha) * T(alpha),
                     f);
  AJ_LAUNCH_CHECK("k_dyn_lm");
  hipLaunchKernelGGL(k_dyn_final<T>, grid, dim3(256), 0, st, nx, ny, T(delta) * T(delta), nu, f,
                     nsgs, ntot, cd);
  AJ_LAUNCH_CHECK("k_dyn_final");
  AJ_HIP(hipFreeAsync(mem, st));
  return AJ_OK;
}

}  // namespace aj

using namespace aj;

extern "C" int aj_les_dynamic_smagorinsky(int dtype, const void* u, const void* v, double nu,
                                          void* nu_sgs, void* nu_total, void* c_d, int nx, int ny,
                                          double dx, double dy, double delta, double alpha,
                                          void* stream) {
  if (!u || !v || !nu_sgs) return fail(AJ_ERR_INVALID, "aj_les_dynamic_smagorinsky: null pointer");
  if (nx < 3 || ny < 3) return fail(AJ_ERR_INVALID, "aj_les_dynamic_smagorinsky: grid must be at least 3x3");
  hipStream_t st = (hipStream_t)stream;
  if (dtype == AJ_F32)
    return dyn_launch<float>((const float*)u, (const float*)v, (float)nu, (float*)nu_sgs,
                             (float*)nu_total, (float*)c_d, nx, ny, dx, dy, delta, alpha, st);
  if (dtype == AJ_F64)
    return dyn_launch<double>((const double*)u, (const double*)v, nu, (double*)nu_sgs,
                              (double*)nu_total, (double*)c_d, nx, ny, dx, dy, delta, alpha, st);
  return fail(AJ_ERR_INVALID, "aj_les_dynamic_smagorinsky: bad dtype");
}
