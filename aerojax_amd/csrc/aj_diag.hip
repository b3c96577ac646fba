// Flow diagnostics of the step, one fused reduction pass:
//   * pressure residual of the viewer / debug path
//     (viewer/simulation_controller.py:107-121, solver/pure_functions/
//     step.py:229-235): ||lap_x(p) - rhs||_2 with rhs = div_nonperiodic(u, v)
//     / dt and lap_x = laplacian_nonperiodic_x (solver/operators.py:67-72:
//     edge-padded in x, periodic roll in y), relative to ||rhs||_2 + 1e-10;
//   * divergence over the pure fluid (simulation_controller.py:186-194):
//     div = np.gradient(u, dx, axis 0) + np.gradient(v, dy, axis 1)
//     (central inside, one-sided first order on the edges), fluid = mask >
//     0.99, div_rms = sqrt(sum (div fluid)^2 / (sum fluid + 1e-8)),
//     l2_div = sqrt(sum (div fluid)^2 dx dy);
//   * check_cfl (advection_schemes/utils.py:14-20): max_speed =
//     max sqrt(u^2 + v^2), cfl = max(max_speed dt / dx, max_speed dt / dy).
// FP64 partial sums (and maxima) per block, fixed-order final reduction.
#include <mutex>
#include <string>

#include "aj_common.h"

namespace aj {

template <typename T>
__global__ __launch_bounds__(256) void k_diag(const T* __restrict__ u, const T* __restrict__ v,
                                              const T* __restrict__ p, const T* __restrict__ mask,
                                              int nx, int ny, double dx, double dy, double dt,
                                              double* __restrict__ part) {
  __shared__ double sh[16];
  const long n = (long)nx * ny;
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  double res2 = 0.0, rhs2 = 0.0, div2 = 0.0, fluid = 0.0, speed = 0.0;
  if (c < n) {
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    // rhs = divergence_nonperiodic(u, v) / dt: zero on the border
    double rhs = 0.0;
    if (i > 0 && i < nx - 1 && j > 0 && j < ny - 1)
      rhs = ((double)u[c + ny] - (double)u[c - ny]) / (2.0 * dx) +
            ((double)v[c + 1] - (double)v[c - 1]) / (2.0 * dy);
    rhs /= dt;
    const double pc = p[c];
    const double pW = p[i > 0 ? c - ny : c], pE = p[i < nx - 1 ? c + ny : c];
    const double pS = p[j > 0 ? c - 1 : c + ny - 1], pN = p[j < ny - 1 ? c + 1 : c - ny + 1];
    const double lap = (pE - 2.0 * pc + pW) / (dx * dx) + (pS + pN - 2.0 * pc) / (dy * dy);
    res2 = (lap - rhs) * (lap - rhs);
    rhs2 = rhs * rhs;
    // np.gradient divergence
    double gx, gy;
    if (i == 0) gx = ((double)u[c + ny] - (double)u[c]) / dx;
    else if (i == nx - 1) gx = ((double)u[c] - (double)u[c - ny]) / dx;
    else gx = ((double)u[c + ny] - (double)u[c - ny]) / (2.0 * dx);
    if (j == 0) gy = ((double)v[c + 1] - (double)v[c]) / dy;
    else if (j == ny - 1) gy = ((double)v[c] - (double)v[c - 1]) / dy;
    else gy = ((double)v[c + 1] - (double)v[c - 1]) / (2.0 * dy);
    if ((double)mask[c] > 0.99) {
      div2 = (gx + gy) * (gx + gy);
      fluid = 1.0;
    }
    const double uc = u[c], vc = v[c];
    speed = sqrt(uc * uc + vc * vc);
  }
  const double a = block_sum(res2, sh);
  const double b = block_sum(rhs2, sh);
  const double d = block_sum(div2, sh);
  const double f = block_sum(fluid, sh);
  const double s = block_max(speed, sh);
  if (threadIdx.x == 0) {
    double* o = part + 5 * (long)blockIdx.x;
    o[0] = a;
    o[1] = b;
    o[2] = d;
    o[3] = f;
    o[4] = s;
  }
}

__global__ __launch_bounds__(256) void k_diag_final(const double* __restrict__ part, int np,
                                                    double dx, double dy, double dt,
                                                    aj_flow_diag* __restrict__ out) {
  __shared__ double sh[16];
  double acc[4] = {0.0, 0.0, 0.0, 0.0}, smax = 0.0;
  for (int k = threadIdx.x; k < np; k += 256) {
    for (int q = 0; q < 4; ++q) acc[q] += part[5 * (long)k + q];
    smax = fmax(smax, part[5 * (long)k + 4]);
  }
  double t[4];
  for (int q = 0; q < 4; ++q) t[q] = block_sum(acc[q], sh);
  const double vmax = block_max(smax, sh);
  if (threadIdx.x == 0) {
    out->max_speed = vmax;
    out->cfl = fmax(vmax * dt / dx, vmax * dt / dy);
    out->pressure_residual = sqrt(t[0]);
    out->rhs_norm = sqrt(t[1]);
    out->relative_residual = sqrt(t[0]) / (sqrt(t[1]) + 1e-10);
    out->div_rms = sqrt(t[2] / (t[3] + 1e-8));
    out->l2_div = sqrt(t[2] * dx * dy);
    out->fluid_cells = t[3];
  }
}

// Launches the two reduction kernels into `part` (5 doubles per 256-cell
// block) and `dres`, copies the result to the host struct and waits.
static int diag_run(int dtype, const void* u, const void* v, const void* p, const void* mask,
                    int nx, int ny, double dx, double dy, double dt, double* part,
                    aj_flow_diag* dres, aj_flow_diag* out, hipStream_t st) {
  const long n = (long)nx * ny;
  const int nb = (int)((n + 255) / 256);
  if (dtype == AJ_F32)
    hipLaunchKernelGGL(k_diag<float>, dim3(nb), dim3(256), 0, st, (const float*)u,
                       (const float*)v, (const float*)p, (const float*)mask, nx, ny, dx, dy, dt,
                       part);
  else
    hipLaunchKernelGGL(k_diag<double>, dim3(nb), dim3(256), 0, st, (const double*)u,
                       (const double*)v, (const double*)p, (const double*)mask, nx, ny, dx, dy,
                       dt, part);
  AJ_LAUNCH_CHECK("k_diag");
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, st, (const double*)part, nb, dx, dy, dt,
                     dres);
  AJ_LAUNCH_CHECK("k_diag_final");
  AJ_HIP(hipMemcpyAsync(out, dres, sizeof(aj_flow_diag), hipMemcpyDeviceToHost, st));
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}

static int diag_check(const char* who, int dtype, const void* u, const void* v, const void* p,
                      const void* mask, int nx, int ny, aj_flow_diag* out) {
  if (!u || !v || !p || !mask || !out) return fail(AJ_ERR_INVALID, std::string(who) + ": null pointer");
  if (nx < 3 || ny < 3) return fail(AJ_ERR_INVALID, std::string(who) + ": grid must be at least 3x3");
  if (dtype != AJ_F32 && dtype != AJ_F64) return fail(AJ_ERR_INVALID, std::string(who) + ": bad dtype");
  return AJ_OK;
}

// ---- allocation-free variant ---------------------------------------------
// The partial sums live in a file-scope device array instead of a
// stream-ordered allocation.  Every block of k_diag writes its 5 partials and
// k_diag_final writes the whole result, so nothing has to start at zero.
constexpr long kDiagNoAllocMaxCells = 1L << 23;  // 8,388,608 cells
constexpr int kDiagNoAllocMaxBlocks = (int)(kDiagNoAllocMaxCells / 256);
constexpr int kDiagResultDoubles = (int)((sizeof(aj_flow_diag) + sizeof(double) - 1) / sizeof(double));
static __device__ __align__(256) double scratch[5 * (long)kDiagNoAllocMaxBlocks + kDiagResultDoubles];

// The array's host-side name is not a device address: each device's copy is
// looked up once with that device current and kept here.
constexpr int kDiagMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kDiagMaxDevices] = {};

static int diag_scratch(double** out) {
  int dev = 0;
  AJ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kDiagMaxDevices)
    return fail(AJ_ERR_UNSUPPORTED, "aj_flow_diagnostics_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    AJ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(AJ_ERR_HIP, "aj_flow_diagnostics_noalloc: no address for the scratch array");
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return AJ_OK;
}

}  // namespace aj

using namespace aj;

extern "C" int aj_flow_diagnostics(int dtype, const void* u, const void* v, const void* p,
                                   const void* mask, int nx, int ny, double dx, double dy,
                                   double dt, aj_flow_diag* out, void* stream) {
  if (int rc = diag_check("aj_flow_diagnostics", dtype, u, v, p, mask, nx, ny, out)) return rc;
  hipStream_t st = (hipStream_t)stream;
  const long n = (long)nx * ny;
  const int nb = (int)((n + 255) / 256);
  void* buf = nullptr;
  AJ_HIP(hipMallocAsync(&buf, (size_t)nb * 5 * sizeof(double) + sizeof(aj_flow_diag), st));
  double* part = (double*)buf;
  aj_flow_diag* dres = (aj_flow_diag*)(part + 5 * (long)nb);
  const int rc = diag_run(dtype, u, v, p, mask, nx, ny, dx, dy, dt, part, dres, out, st);
  const hipError_t e = hipFreeAsync(buf, st);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "hipFreeAsync");
  return AJ_OK;
}

extern "C" int aj_flow_diagnostics_noalloc(int dtype, const void* u, const void* v, const void* p,
                                           const void* mask, int nx, int ny, double dx, double dy,
                                           double dt, aj_flow_diag* out, void* stream) {
  if (int rc = diag_check("aj_flow_diagnostics_noalloc", dtype, u, v, p, mask, nx, ny, out))
    return rc;
  if ((long)nx * ny > kDiagNoAllocMaxCells)
    return fail(AJ_ERR_INVALID, "aj_flow_diagnostics_noalloc: nx*ny = " +
                                    std::to_string((long)nx * ny) + " exceeds the maximum of " +
                                    std::to_string(kDiagNoAllocMaxCells) + " cells");
  double* part = nullptr;
  if (int rc = diag_scratch(&part)) return rc;
  aj_flow_diag* dres = (aj_flow_diag*)(part + 5 * (long)kDiagNoAllocMaxBlocks);
  return diag_run(dtype, u, v, p, mask, nx, ny, dx, dy, dt, part, dres, out, (hipStream_t)stream);
}
