// Obstacle geometry on the device: the level set (SDF) of a cylinder or a
// NACA 4-digit section and the smoothed fluid mask, evaluated per cell on the
// reference's grid coordinates.  Shared by the single-grid entry point and
// the batched driver.  All arithmetic in FP64 (the reference's host setup
// precision), stored in the field type.
#pragma once

#include "aj_common.h"

namespace aj {

// numpy.linspace(0, l, n)[i]: i * (l / (n - 1)), the last point exactly l
// (GridParams.__post_init__, solver/params.py:54-56)
__device__ __forceinline__ double linspace_at(int i, int n, double l) {
  if (n < 2) return 0.0;
  return (i == n - 1) ? l : (double)i * (l / (double)(n - 1));
}

// np.minimum: NaN if either operand is NaN
__device__ __forceinline__ double nan_min(double a, double b) {
  return (isnan(a) || isnan(b)) ? (a + b) : (a < b ? a : b);
}

// SDF at (X, Y):
//  cylinder: sqrt((X - cx)^2 + (Y - cy)^2) - r
//    (solver/simulation/mask_generator.py:151-160, solver/geometry.py:10-13)
//  NACA: naca_surface_distance (obstacles/naca_airfoils.py:109-178):
//    rotate about the leading edge by +AoA, vertical distance to the upper /
//    lower surface inside the chord, distance to the LE / TE point outside,
//    "inside" with a 0.01 (chord units) tolerance; a symmetric section
//    (m = p = 0) keeps the reference's 0/0 camber terms, so x_norm == 0
//    gives NaN as there.
__device__ __forceinline__ double obstacle_sdf(const aj_obstacle& o, double ca, double sa,
                                               double X, double Y) {
  if (o.kind == AJ_OBSTACLE_CYLINDER) {
    const double ex = X - o.x0, ey = Y - o.y0;
    return sqrt(ex * ex + ey * ey) - o.radius;
  }
  const double xl = X - o.x0, yl = Y - o.y0;
  const double xr = xl * ca - yl * sa;
  const double yr = xl * sa + yl * ca;
  const double c = o.chord;
  const double xn = xr / c, yn = yr / c;
  const bool in_chord = (xn >= 0.0) && (xn <= 1.0);
  const double xs = fmin(fmax(xn, 0.0), 1.0);
  const double xs2 = xs * xs, xs3 = xs2 * xs, xs4 = xs3 * xs;
  const double yt =
      5.0 * o.t * (0.2969 * sqrt(fabs(xs)) - 0.1260 * xs - 0.3516 * xs2 + 0.2843 * xs3 - 0.1015 * xs4);
  const double m = o.m, p = o.p;
  double yc, dyc;
  if (xn <= p) {
    yc = m / (p * p) * (2.0 * p * xn - xn * xn);
    dyc = 2.0 * m / (p * p) * (p - xn);
  } else {
    const double q = (1.0 - p) * (1.0 - p);
    yc = m / q * ((1.0 - 2.0 * p) + 2.0 * p * xn - xn * xn);
    dyc = 2.0 * m / q * (p - xn);
  }
  const double ct = cos(atan(dyc));
  const double yu = yc + yt * ct, ylo = yc - yt * ct;
  const double d_surf = nan_min(fabs(yn - yu) * c, fabs(yn - ylo) * c);
  const double d_le = sqrt(xr * xr + yr * yr);
  const double d_te = sqrt((xr - c) * (xr - c) + yr * yr);
  const double dist = in_chord ? d_surf : nan_min(d_le, d_te);
  const bool inside = in_chord && (yn >= ylo - 0.01) && (yn <= yu + 0.01);
  return inside ? -dist : dist;
}

// mask = 1 - sigmoid(-sdf / eps) (mask_generator.py:71-72, 157-158)
__device__ __forceinline__ double obstacle_mask(double sdf, double eps) {
  const double z = -sdf / eps;
  return 1.0 - 1.0 / (1.0 + exp(-z));
}

}  // namespace aj
