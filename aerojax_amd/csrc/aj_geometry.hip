// Single-grid obstacle geometry: SDF and mask of one obstacle on an nx x ny
// grid of the reference's coordinates (aj_geometry).
#include <cmath>

#include "aj_geometry_kernels.h"

namespace aj {

template <typename T>
__global__ __launch_bounds__(256) void k_geometry(aj_obstacle o, double ca, double sa, int nx,
                                                  int ny, double lx, double ly, T* __restrict__ sdf,
                                                  T* __restrict__ mask) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= ny) return;
  const double s = obstacle_sdf(o, ca, sa, linspace_at(i, nx, lx), linspace_at(j, ny, ly));
  const long c = (long)i * ny + j;
  if (sdf) sdf[c] = (T)s;
  if (mask) mask[c] = (T)obstacle_mask(s, o.eps);
}

// the rotation of the NACA frame, computed once on the host as the
// reference does (jnp.radians, jnp.cos / jnp.sin of a scalar)
void obstacle_rotation(const aj_obstacle& o, double* ca, double* sa) {
  const double a = o.aoa_deg * (3.14159265358979323846 / 180.0);
  *ca = std::cos(a);
  *sa = std::sin(a);
}

int check_obstacle(const aj_obstacle& o, const char* who) {
  char msg[160];
  const bool ok = (o.kind == AJ_OBSTACLE_CYLINDER && o.radius > 0 && o.eps > 0) ||
                  (o.kind == AJ_OBSTACLE_NACA && o.chord > 0 && o.eps > 0);
  if (ok) return AJ_OK;
  std::snprintf(msg, sizeof(msg), "%s: bad obstacle (kind %d, radius/chord > 0, eps > 0)", who,
                o.kind);
  return fail(AJ_ERR_INVALID, msg);
}

}  // namespace aj

using namespace aj;

extern "C" int aj_geometry(int dtype, const aj_obstacle* obs, int nx, int ny, double lx,
                           double ly, void* sdf, void* mask, void* stream) {
  if (!obs || (!sdf && !mask)) return fail(AJ_ERR_INVALID, "aj_geometry: null pointer");
  if (nx < 2 || ny < 2 || !(lx > 0) || !(ly > 0)) return fail(AJ_ERR_INVALID, "aj_geometry: bad grid");
  if (int rc = check_obstacle(*obs, "aj_geometry")) return rc;
  double ca, sa;
  obstacle_rotation(*obs, &ca, &sa);
  hipStream_t st = (hipStream_t)stream;
  const dim3 grid(cdiv(ny, 256), nx);
  if (dtype == AJ_F32)
    hipLaunchKernelGGL(k_geometry<float>, grid, dim3(256), 0, st, *obs, ca, sa, nx, ny, lx, ly,
                       (float*)sdf, (float*)mask);
  else if (dtype == AJ_F64)
    hipLaunchKernelGGL(k_geometry<double>, grid, dim3(256), 0, st, *obs, ca, sa, nx, ny, lx, ly,
                       (double*)sdf, (double*)mask);
  else
    return fail(AJ_ERR_INVALID, "aj_geometry: bad dtype");
  AJ_LAUNCH_CHECK("k_geometry");
  return AJ_OK;
}
