// Batched stepping of many independent simulations (each with its own grid
// shape, spacing, geometry and viscosity) in shared launches.
//
// Instances are grouped by grid shape; every field of the batch is one
// buffer in which the instances follow each other group by group (instance
// k occupies [off_k, off_k + nx_g * ny_g)).  Each kernel launch carries a
// BatchWork descriptor per group (shape, spacing, first instance, element
// base, work items per instance) and runs as a persistent grid: workgroup w
// processes work items w, w + gridDim.x, ... and maps each to (group,
// instance, tile) by a scan of the <= kMaxGroups group starts.  Per-instance
// scalars (viscosity) and the per-instance multigrid gates live in device
// arrays indexed by the instance id.
//
// One batched step is step_pure's composition (solver/pure_functions/
// step.py:49-467, as aj_rk_step + aj_divergence_rhs + aj_poisson_mg
// (reference) + aj_project) for every instance: [Smagorinsky rows], RK3
// predictor tiles, rhs rows, reference multigrid V-cycles level by level
// (all instances of a level in one launch; the 10-sweep coarsest solve one
// workgroup per instance) with a per-instance convergence gate, projection
// rows.  The launch sequence of a step has no host synchronisation and is
// captured once into a hipGraph per set of buffers.
//
// With the textbook solver the pressure of every instance is solved to the
// tolerance instead: one batched textbook plan per group (its instances in
// blockIdx.z, each with its own gate, guess test and live buffer), all
// groups' outer iterations enqueued together and checked with one host
// synchronisation per round.
#include <cstdio>
#include <cstring>
#include <type_traits>
#include <vector>

#include "aj_geometry_kernels.h"
#include "aj_mg.h"
#include "aj_mg_ref_kernels.h"
#include "aj_step_kernels.h"
#include "aj_rk_rc.h"

namespace aj {

constexpr int kMaxGroups = 8;
constexpr int kBatchRkNTH = 512;
template <typename T> struct BatchRkTile;  // RK tile: 16x64 (FP32) / 8x64 (FP64) keeps LDS < 64 KiB
template <> struct BatchRkTile<float> { static constexpr int TX = 16, TY = 64; };
template <> struct BatchRkTile<double> { static constexpr int TX = 8, TY = 64; };
constexpr int kBatchJX = 16, kBatchJY = 64;  // Jacobi tile (rows of >= 64 cells)
constexpr int kResChunks = 16;  // residual-norm partial sums per instance
constexpr int kWgCoarse = 2;    // coarsest solve: 1024-thread workgroups per CU

struct BatchWork {
  int ng, total;
  int wstart[kMaxGroups + 1];  // work-item prefix over groups
  int per_inst[kMaxGroups];    // work items per instance
  int first[kMaxGroups];       // first instance id of the group
  int nx[kMaxGroups], ny[kMaxGroups];
  int tiles_y[kMaxGroups];     // tile kernels: tiles along y
  long base[kMaxGroups];       // element offset of the group's first instance
  long cbase[kMaxGroups];      // coarse-level buffers: same for the next level
  int cnx[kMaxGroups], cny[kMaxGroups];
  double dx[kMaxGroups], dy[kMaxGroups];
};

struct BatchItem {
  int g, inst, sub;
  long off, coff;
};

// Persistent-grid order of the work items: workgroups go to the 8 XCDs
// round-robin, so with a grid that is a multiple of 8 the workgroups of XCD x
// take the contiguous eighth x of the items between them.  Neighbouring tiles
// and rows (which share halo rows and cache lines) are then read through the
// same L2 instead of one per XCD.  Loop: for (XcdItems q(total); q.t < q.end;
// q.t += q.step) { item = q.base + q.t; ... }
struct XcdItems {
  int base, t, step, end;
  __device__ __forceinline__ explicit XcdItems(int total) {
    if ((gridDim.x & 7) == 0) {
      const int per = (total + 7) >> 3;
      base = (int)(blockIdx.x & 7) * per;
      t = (int)(blockIdx.x >> 3);
      step = (int)(gridDim.x >> 3);
      end = min(per, total - base);
    } else {
      base = 0;
      t = (int)blockIdx.x;
      step = (int)gridDim.x;
      end = total;
    }
  }
};

__device__ __forceinline__ BatchItem batch_item(const BatchWork& w, int item) {
  int g = 0;
#pragma unroll
  for (int k = 1; k < kMaxGroups; ++k)
    if (k < w.ng && item >= w.wstart[k]) g = k;
  const int local = item - w.wstart[g];
  const int k = local / w.per_inst[g];
  BatchItem it;
  it.g = g;
  it.inst = w.first[g] + k;
  it.sub = local - k * w.per_inst[g];
  it.off = w.base[g] + (long)k * w.nx[g] * w.ny[g];
  it.coff = w.cbase[g] + (long)k * w.cnx[g] * w.cny[g];
  return it;
}

// ---------------------------------------------------------------- step kernels

template <typename T>
struct BatchStepArgs {
  const T *u, *v, *mask, *sdf;
  T *us, *vs, *rhs, *p, *uo, *vo;
  T* nsgs;           // Smagorinsky nu_sgs, or null (no LES)
  const double* nu;  // [instances]
  double dt, U_inf, inv_eta, cs;
};

template <typename T, int TX, int TY, int NTH>
__global__ __launch_bounds__(NTH) void kb_rk(BatchWork w, BatchStepArgs<T> s) {
  extern __shared__ unsigned char smem_raw[];
  T* sm = reinterpret_cast<T*>(smem_raw);
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g;
    RkArgs<T> a;
    a.u = s.u + it.off;
    a.v = s.v + it.off;
    a.mask = s.mask + it.off;
    a.sdf = s.sdf ? s.sdf + it.off : nullptr;
    a.nu_sgs = s.nsgs ? s.nsgs + it.off : nullptr;
    a.nu_field = nullptr;
    a.u_out = s.us + it.off;
    a.v_out = s.vs + it.off;
    a.rhs = nullptr;
    a.rhs_masked = 0;
    a.nx = w.nx[g];
    a.ny = w.ny[g];
    a.dt = T(s.dt);
    a.nu = T(s.nu[it.inst]);
    a.inv_dx = T(1.0 / w.dx[g]);
    a.inv_dy = T(1.0 / w.dy[g]);
    a.inv_dx2 = T(1.0 / (w.dx[g] * w.dx[g]));
    a.inv_dy2 = T(1.0 / (w.dy[g] * w.dy[g]));
    a.U_inf = T(s.U_inf);
    a.inv_eta = T(s.inv_eta);
    a.sdf_scale = T(1.0 / (0.1 * w.dx[g]));
    a.hyper = T(0);
    a.slip = 1;
    const int ty = it.sub % w.tiles_y[g], tx = it.sub / w.tiles_y[g];
    rk_tile<T, TX, TY, false, false, NTH>(a, tx * TX, ty * TY, sm);
    __syncthreads();  // LDS reused by the next item
  }
}

// register-column RK3 predictor with the pressure rhs in its epilogue (FP32,
// every group with even ny >= 4): items are tiles of one instance, x-tiles
// of the whole row width when ny <= 128, else x-tiles x column strips
// (w.tiles_y strips per x-tile)
template <int W, int R>
__global__ __launch_bounds__(W * 64) void kb_rk_rc(BatchWork w, BatchStepArgs<float> s) {
  using G = RkRcGeo<W, R, false, true>;
  __shared__ float4 xb[2][W][2][64];
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const BatchItem it = batch_item(w, q.base + q.t);
    const int g = it.g;
    RkRcArgs a;
    a.u = s.u + it.off;
    a.v = s.v + it.off;
    a.mask = s.mask + it.off;
    a.sdf = s.sdf ? s.sdf + it.off : nullptr;
    a.nu_sgs = s.nsgs ? s.nsgs + it.off : nullptr;
    a.nu_field = nullptr;
    a.u_out = s.us + it.off;
    a.v_out = s.vs + it.off;
    a.rhs = s.rhs + it.off;
    a.rhs_masked = 1;
    a.nx = w.nx[g];
    a.ny = w.ny[g];
    a.dt = (float)s.dt;
    a.nu = (float)s.nu[it.inst];
    a.inv_dx = (float)(1.0 / w.dx[g]);
    a.inv_dy = (float)(1.0 / w.dy[g]);
    a.inv_dx2 = (float)(1.0 / (w.dx[g] * w.dx[g]));
    a.inv_dy2 = (float)(1.0 / (w.dy[g] * w.dy[g]));
    a.U_inf = (float)s.U_inf;
    a.inv_eta = (float)s.inv_eta;
    a.sdf_scale = (float)(1.0 / (0.1 * w.dx[g]));
    a.inv_2dx = (float)(1.0 / (2.0 * w.dx[g]));
    a.inv_2dy = (float)(1.0 / (2.0 * w.dy[g]));
    a.inv_dt = (float)(1.0 / s.dt);
    a.slip = 1;
    const int ty = it.sub % w.tiles_y[g], tx = it.sub / w.tiles_y[g];
    int gx0, kx0, kx1;
    G::x_tile(tx, a.nx, &gx0, &kx0, &kx1);
    if (a.ny <= 128) {
      rk_rc_tile<W, R, false, true>(a, gx0, kx0, kx1, 0, 0, a.ny, xb);
    } else {
      const int gy0 = ty * G::UY - G::HY;
      rk_rc_tile<W, R, false, true>(a, gx0, kx0, kx1, gy0, gy0 + G::HY, gy0 + G::HY + G::UY, xb);
    }
    __syncthreads();  // the exchange buffers are reused by the next item
  }
}

// cell items: kCellsPerItem consecutive cells (row-major) of one instance,
// 4 per thread, so that every lane has a cell whatever the row length (a row
// of 128 cells would leave half of a 256-thread workgroup idle) and the four
// cells' loads are in flight together; a thread steps its (row, column) by
// 256 cells.  f(i, j) runs for each cell of item `sub` of an nx x ny grid.
constexpr int kCellsPerItem = 1024;
template <typename F>
__device__ __forceinline__ void for_item_cells(int sub, int nx, int ny, F&& f) {
  const int c0 = sub * kCellsPerItem + (int)threadIdx.x;
  const int di = 256 / ny, dj = 256 - di * ny;
  int i = c0 / ny, j = c0 - (c0 / ny) * ny;
#pragma unroll
  for (int k = 0; k < kCellsPerItem / 256; ++k) {
    if (i < nx) f(i, j);
    i += di;
    j += dj;
    if (j >= ny) {
      j -= ny;
      ++i;
    }
  }
}

// nu_sgs = constant Smagorinsky with delta = sqrt(dx dy) (step.py:111-122)
template <typename T>
__global__ __launch_bounds__(256) void kb_les(BatchWork w, BatchStepArgs<T> s) {
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    const double cd = s.cs * sqrt(w.dx[g] * w.dy[g]);
    const T coef = T(cd * cd);
    for_item_cells(it.sub, nx, ny, [&](int i, int j) {
      s.nsgs[it.off + (long)i * ny + j] =
          les_cell(s.u + it.off, s.v + it.off, nx, ny, i, j, T(1.0 / (2.0 * w.dx[g])),
                   T(1.0 / (2.0 * w.dy[g])), coef);
    });
  }
}

// obstacle geometry per instance (rotation precomputed on the host)
struct ObstacleDev {
  aj_obstacle o;
  double ca, sa;
};

template <typename T>
__global__ __launch_bounds__(256) void kb_geometry(BatchWork w, const ObstacleDev* __restrict__ obs,
                                                   T* __restrict__ sdf, T* __restrict__ mask) {
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    const ObstacleDev od = obs[it.inst];
    const double lx = nx * w.dx[g], ly = ny * w.dy[g];
    for_item_cells(it.sub, nx, ny, [&](int i, int j) {
      const double s = obstacle_sdf(od.o, od.ca, od.sa, linspace_at(i, nx, lx), linspace_at(j, ny, ly));
      const long c = it.off + (long)i * ny + j;
      if (sdf) sdf[c] = (T)s;
      if (mask) mask[c] = (T)obstacle_mask(s, od.o.eps);
    });
  }
}

// rhs = div * mask / dt
template <typename T>
__global__ __launch_bounds__(256) void kb_div(BatchWork w, BatchStepArgs<T> s) {
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    for_item_cells(it.sub, nx, ny, [&](int i, int j) {
      div_rhs_cell(s.us + it.off, s.vs + it.off, s.mask + it.off, s.rhs + it.off, nx, ny, i, j,
                   T(1.0 / (2.0 * w.dx[g])), T(1.0 / (2.0 * w.dy[g])), T(1.0 / s.dt));
    });
  }
}

// for_item_cells over the pairs (i, j), (i, j + 1) of even j (ny even): 2
// pairs per thread, f(i, j) runs for the first cell of each pair
template <typename F>
__device__ __forceinline__ void for_item_pairs(int sub, int nx, int ny, F&& f) {
  const int hy = ny >> 1;
  const int q0 = sub * (kCellsPerItem / 2) + (int)threadIdx.x;
  const int di = 256 / hy, dq = 256 - di * hy;
  int i = q0 / hy, q = q0 - (q0 / hy) * hy;
#pragma unroll
  for (int k = 0; k < kCellsPerItem / 512; ++k) {
    if (i < nx) f(i, 2 * q);
    i += di;
    q += dq;
    if (q >= hy) {
      q -= hy;
      ++i;
    }
  }
}

// pairs: the six arrays are pair-aligned; an instance goes by pairs when its
// row length is even and its offset keeps that alignment
template <typename T>
__global__ __launch_bounds__(256) void kb_project(BatchWork w, BatchStepArgs<T> s, int pairs) {
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    const T dt = T(s.dt), i2x = T(1.0 / (2.0 * w.dx[g])), i2y = T(1.0 / (2.0 * w.dy[g]));
    const T thr = T(0.1 / w.dx[g]), U = T(s.U_inf);
    const T *us = s.us + it.off, *vs = s.vs + it.off, *p = s.p + it.off, *m = s.mask + it.off;
    T *uo = s.uo + it.off, *vo = s.vo + it.off;
    if (pairs && !((ny | (int)it.off) & 1)) {
      for_item_pairs(it.sub, nx, ny, [&](int i, int j) {
        project_pair(us, vs, p, m, uo, vo, nx, ny, i, j, dt, i2x, i2y, thr, U);
      });
    } else {
      for_item_cells(it.sub, nx, ny, [&](int i, int j) {
        project_cell(us, vs, p, m, uo, vo, nx, ny, i, j, dt, i2x, i2y, thr, U);
      });
    }
  }
}

// ---------------------------------------------------------------- multigrid

template <typename T>
struct BatchMgArgs {
  const T* p_in;  // null -> zero guess
  const T* e_c;   // coarse correction (level + 1 buffers), or null
  const T* b;
  T* p_out;
  int bzl, clamp;
  aj_mg_stats* stats;  // [instances]
  // non-null: each tile item also writes the sums of the squared residual
  // of its input iterate (one per wave) to
  // res_partial[instance * res_stride + 4 tile + wave]
  double* res_partial;
  int res_stride;
};

template <typename T, int TX, int TY>
__global__ __launch_bounds__(256) void kb_jacobi(BatchWork w, BatchMgArgs<T> m) {
  __shared__ JacobiLds<T, TX, TY, 2> L;
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g;
    RefJacobiArgs<T> a;
    a.p_in = m.p_in ? m.p_in + it.off : nullptr;
    a.e_c = m.e_c ? m.e_c + it.coff : nullptr;
    a.ncx = w.cnx[g];
    a.ncy = w.cny[g];
    a.b = m.b + it.off;
    a.b_zero_last = m.bzl;
    a.p_out = m.p_out + it.off;
    a.nx = w.nx[g];
    a.ny = w.ny[g];
    // fine-grid spacing on every level (multigrid_solver.py:158-171)
    a.ax = T(1) / (T(w.dx[g]) * T(w.dx[g]));
    a.ay = T(1) / (T(w.dy[g]) * T(w.dy[g]));
    a.inv_diag = T(1) / (T(2) * (a.ax + a.ay));
    a.clamp_last = m.clamp;
    a.gate = m.stats + it.inst;
    a.res_out = m.res_partial ? m.res_partial + (long)it.inst * m.res_stride + 4 * it.sub : nullptr;
    const int ty = it.sub % w.tiles_y[g], tx = it.sub / w.tiles_y[g];
    ref_jacobi_tile<T, TX, TY, 2>(a, tx * TX, ty * TY, L);
    __syncthreads();
  }
}

// residual + restriction: cell items of the coarse grid
template <typename T>
__global__ __launch_bounds__(256) void kb_resrestrict(BatchWork w, BatchMgArgs<T> m, T* bc) {
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g;
    if (gated(m.stats + it.inst)) continue;
    const int ncx = w.cnx[g], ncy = w.cny[g];
    const T ax = T(1) / (T(w.dx[g]) * T(w.dx[g])), ay = T(1) / (T(w.dy[g]) * T(w.dy[g]));
    const int degenerate = (w.nx[g] / 2 < 2 || w.ny[g] / 2 < 2) ? 1 : 0;
    for_item_cells(it.sub, ncx, ncy, [&](int I, int J) {
      ref_resrestrict_cell(m.p_out + it.off, m.b + it.off, m.bzl, w.nx[g], w.ny[g], ax, ay,
                           bc + it.coff, ncy, degenerate, I, J);
    });
  }
}

// coarsest level: 10 sweeps from zero, one workgroup per instance
template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void kb_coarse(BatchWork w, BatchMgArgs<T> m) {
  extern __shared__ unsigned char smem_raw[];
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g;
    if (gated(m.stats + it.inst)) continue;
    const T ax = T(1) / (T(w.dx[g]) * T(w.dx[g])), ay = T(1) / (T(w.dy[g]) * T(w.dy[g]));
    ref_coarse_wg(m.b + it.off, m.p_out + it.off, w.nx[g], w.ny[g], ax, ay,
                  T(1) / (T(2) * (ax + ay)), 10, reinterpret_cast<T*>(smem_raw));
    __syncthreads();
  }
}

// levels sub_lev .. coarsest of every group, for the in-LDS V-cycle
struct SubDims {
  int nl;
  int nx[kMaxGroups][kMaxRefLevels], ny[kMaxGroups][kMaxRefLevels];
};

// the V-cycle from level sub_lev down for one instance per workgroup, all
// in LDS (ref_subcycle_wg), in place of the per-level launches
template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void kb_subcycle(BatchWork w, SubDims d,
                                                              BatchMgArgs<T> m) {
  extern __shared__ unsigned char smem_raw[];
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g;
    if (gated(m.stats + it.inst)) continue;
    const T ax = T(1) / (T(w.dx[g]) * T(w.dx[g])), ay = T(1) / (T(w.dy[g]) * T(w.dy[g]));
    ref_subcycle_wg(m.b + it.off, m.p_out + it.off, d.nx[g], d.ny[g], d.nl, ax, ay,
                    T(1) / (T(2) * (ax + ay)), reinterpret_cast<T*>(smem_raw));
    __syncthreads();  // LDS reused by the next instance
  }
}

// sum(r^2) partials: kResChunks items per instance, item k taking the
// contiguous rows [k nx / kResChunks, (k + 1) nx / kResChunks), written to
// partial[instance * stride + k]; a thread steps its (row, column) by 256
// cells without a division in the loop
template <typename T>
__global__ __launch_bounds__(256) void kb_resnorm(BatchWork w, const T* __restrict__ p,
                                                  const T* __restrict__ b,
                                                  double* __restrict__ partial, int stride) {
  __shared__ double sh[16];
  for (XcdItems q(w.total); q.t < q.end; q.t += q.step) {
    const int item = q.base + q.t;
    const BatchItem it = batch_item(w, item);
    const int g = it.g, nx = w.nx[g], ny = w.ny[g];
    const int r0 = (int)((long)it.sub * nx / kResChunks);
    const int r1 = (int)((long)(it.sub + 1) * nx / kResChunks);
    const T ax = T(1) / (T(w.dx[g]) * T(w.dx[g])), ay = T(1) / (T(w.dy[g]) * T(w.dy[g]));
    const int di = 256 / ny, dj = 256 - di * ny;
    int i = r0 + (int)threadIdx.x / ny, j = (int)threadIdx.x % ny;
    double acc = 0.0;
#pragma unroll 4
    for (; i < r1;) {
      const double r = (double)ref_residual(p + it.off, b + it.off, 1, nx, ny, ax, ay, i, j);
      acc += r * r;
      i += di;
      j += dj;
      if (j >= ny) {
        j -= ny;
        ++i;
      }
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partial[(long)it.inst * stride + it.sub] = t;
  }
}

// per instance (one wave each): sum its w.per_inst * per_item partials
// (layout w, partial[instance * stride + k]) in a fixed order, record the
// norm, set the gate (multigrid_solver.py:222-234).  An instance whose gate
// was already set ran no items: its iterate is unchanged and keeps its norm.
__global__ __launch_bounds__(64) void kb_resnorm_final(const double* __restrict__ partial,
                                                       BatchWork w, int per_item, int stride,
                                                       const double* __restrict__ dxdy,
                                                       const double* __restrict__ tol2,
                                                       aj_mg_stats* __restrict__ stats) {
  const int inst = blockIdx.x;
  int g = 0;
  for (int k = 1; k < w.ng; ++k)
    if (inst >= w.first[k]) g = k;
  const int n = w.per_inst[g] * per_item;
  double acc = 0.0;
  for (int k = threadIdx.x; k < n; k += 64) acc += partial[(long)inst * stride + k];
  const double tot = wave_sum(acc);
  if (threadIdx.x != 0) return;
  aj_mg_stats* st = stats + inst;
  const int executed = st->converged == 0;
  const double nsq = executed ? tot * dxdy[inst] : st->norm;
  const int k = st->outer;
  if (k < AJ_MG_MAX_NORMS) st->norms[k] = nsq;
  st->outer = k + 1;
  if (executed) st->cycles += 1;
  st->norm = nsq;
  if (!isfinite(nsq)) st->nonfinite = 1;
  st->converged = (nsq < tol2[inst]) ? 1 : 0;
}

__global__ __launch_bounds__(64) void kb_mg_begin(aj_mg_stats* stats, int ninst) {
  const int inst = blockIdx.x;
  if (inst >= ninst) return;
  aj_mg_stats* st = stats + inst;
  if (threadIdx.x == 0) {
    st->cycles = 0;
    st->outer = 0;
    st->converged = 0;
    st->nonfinite = 0;
    st->norm = 0.0;
  }
  st->norms[threadIdx.x] = 0.0;  // 64 threads = AJ_MG_MAX_NORMS
}

}  // namespace aj

#include "aj_batch_rc.h"

using namespace aj;

// ---------------------------------------------------------------- host side

struct aj_batch {
  int dtype = AJ_F32;
  int ng = 0, ninst = 0, levels = 4, ml = 0;
  std::vector<int> gnx, gny, gcount, gfirst;
  std::vector<double> gdx, gdy;
  std::vector<std::vector<int>> lnx, lny;  // [group][level]
  std::vector<long> level_len;             // elements of one field at each level
  std::vector<std::vector<long>> lbase;    // [level][group] element base
  long cells = 0;
  void* ws = nullptr;
  std::vector<void*> B, Y, E;  // per level >= 1 (index level); level 0 uses step buffers
  void *us = nullptr, *vs = nullptr, *rhs = nullptr, *nsgs = nullptr;
  std::vector<aj_mg_plan*> tb;  // textbook solver: one batched plan per group (created on first use)
  int last_solver = AJ_MG_REFERENCE;
  // levels sub_lev .. ml run as one in-LDS V-cycle per instance (0: none)
  int sub_lev = 0;
  size_t sub_lds = 0;
  double* partial = nullptr;  // [instances][res_stride] residual-norm partial sums
  int res_stride = 0;
  double* dxdy = nullptr;
  double* tol2 = nullptr;
  double* nu = nullptr;
  void* obs = nullptr;  // ObstacleDev[instances]
  aj_mg_stats* stats = nullptr;
  double last_tol = -1.0;
  int use_graph = 1;
  int tuned = 0;
  // two cached step graphs (steppers alternate two sets of buffers)
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  const void* key[2][8] = {};
  double key_prm[2][7] = {};
  int victim = 0;
  int device = 0, cus = 256;
};

namespace {

// kind: 0 = tiles tx x ty, 1 = cell items (this level), 2 = cell items of
// the next coarser level, 3 = one item per instance, 4 = kResChunks per
// instance
BatchWork make_work(const aj_batch* b, int level, int kind, int tx = 16, int ty = 64) {
  BatchWork w;
  std::memset(&w, 0, sizeof(w));
  int tot = 0, ng = 0;
  for (int g = 0; g < b->ng; ++g) {
    const int nx = b->lnx[g][level], ny = b->lny[g][level];
    int per = 0;
    w.tiles_y[ng] = (int)cdiv(ny, ty);
    switch (kind) {
      case 0: per = (int)(cdiv(nx, tx) * cdiv(ny, ty)); break;
      case 1: per = (int)cdiv((long)nx * ny, kCellsPerItem); break;
      case 2:
        per = (int)cdiv((long)b->lnx[g][level + 1] * b->lny[g][level + 1], kCellsPerItem);
        break;
      case 3: per = 1; break;
      default: per = kResChunks; break;
    }
    w.wstart[ng] = tot;
    w.per_inst[ng] = per;
    w.first[ng] = b->gfirst[g];
    w.nx[ng] = nx;
    w.ny[ng] = ny;
    w.base[ng] = b->lbase[level][g];
    if (level + 1 <= b->ml) {
      w.cnx[ng] = b->lnx[g][level + 1];
      w.cny[ng] = b->lny[g][level + 1];
      w.cbase[ng] = b->lbase[level + 1][g];
    }
    w.dx[ng] = b->gdx[g];
    w.dy[ng] = b->gdy[g];
    tot += per * b->gcount[g];
    ++ng;
  }
  w.ng = ng;
  w.wstart[ng] = tot;
  w.total = tot;
  return w;
}

unsigned persistent_grid(const aj_batch* b, int total, int per_cu) {
  const long cap = (long)b->cus * per_cu;
  return (unsigned)std::max<long>(1, std::min<long>(total, cap));
}

// persistent grid size, tuned per kernel and work size among
// kWgCands workgroups per CU (outputs never feed the same launch)
constexpr int kWgCands[4] = {4, 8, 16, 32};

template <typename F>
int batch_launch(const aj_batch* b, const char* name, int total, F&& launch, hipStream_t st) {
  char key[96];
  std::snprintf(key, sizeof(key), "batch/%s/%d", name, total);
  auto L = [&](int c, hipStream_t s) -> int {
    return launch(dim3(persistent_grid(b, total, kWgCands[c])), s);
  };
  int pick = 0;
  return tune_pick(key, 4, L, st, &pick);
}

template <int TX_, int TY_>
struct JacobiTile {
  static constexpr int TX = TX_, TY = TY_;
};

// Jacobi launch of one level: 16x64 tiles while every instance's row length
// at the level is at least 64, else 32x32, else 64x16 (a tile wider than the
// grid) and the tile layout in *wt_out when given
template <typename T>
int batch_jacobi(aj_batch* b, int lev, const char* name, const BatchMgArgs<T>& m, hipStream_t st,
                 BatchWork* wt_out = nullptr) {
  int nymin = b->lny[0][lev];
  for (int g = 1; g < b->ng; ++g) nymin = std::min(nymin, b->lny[g][lev]);
  auto go = [&](auto tile) -> int {
    constexpr int TX = decltype(tile)::TX, TY = decltype(tile)::TY;
    const BatchWork wt = make_work(b, lev, 0, TX, TY);
    if (wt_out) *wt_out = wt;
    return batch_launch(b, name, wt.total, [&](dim3 g, hipStream_t s) -> int {
      hipLaunchKernelGGL((kb_jacobi<T, TX, TY>), g, dim3(256), 0, s, wt, m);
      AJ_LAUNCH_CHECK("kb_jacobi");
      return AJ_OK;
    }, st);
  };
  if (nymin >= 64) return go(JacobiTile<kBatchJX, kBatchJY>{});
  if (nymin >= 32) return go(JacobiTile<32, 32>{});
  return go(JacobiTile<64, 16>{});
}

// register-column Jacobi tiles (FP32): waves per workgroup, rows per wave
constexpr int kJacRcW = 8, kJacRcR = 4;
// register-column RK predictor tiles (FP32)
constexpr int kRkRcW = 8, kRkRcR = 4;

// the register-column Jacobi kernels apply on level lev when every group
// has even nx, ny with ny <= 128 (a row is at most one wave of two-column
// lanes) and the next level is the half-size grid with both sides >= 2
bool jac_rc_ok(const aj_batch* b, int lev) {
  if (b->dtype != AJ_F32 || lev + 1 > b->ml) return false;
  for (int g = 0; g < b->ng; ++g) {
    const int nx = b->lnx[g][lev], ny = b->lny[g][lev];
    if ((nx | ny) & 1 || ny > 128 || nx < 4 || ny < 4) return false;
    if (b->lnx[g][lev + 1] != nx / 2 || b->lny[g][lev + 1] != ny / 2) return false;
  }
  return true;
}

// register-column Jacobi launch of one level (PRE: with the residual and
// its restriction into bc); the item layout in *wt_out when given
template <bool PRE>
int batch_jacobi_rc(aj_batch* b, int lev, const char* name, const BatchMgArgs<float>& m,
                    float* bc, hipStream_t st, BatchWork* wt_out = nullptr) {
  using G = JacRcGeo<kJacRcW, kJacRcR, PRE>;
  const BatchWork wt = make_work(b, lev, 0, G::U, 128);
  if (wt_out) *wt_out = wt;
  return batch_launch(b, name, wt.total, [&](dim3 g, hipStream_t s) -> int {
    hipLaunchKernelGGL((kb_jacobi_rc<kJacRcW, kJacRcR, PRE>), g, dim3(kJacRcW * 64), 0, s, wt, m,
                       bc);
    AJ_LAUNCH_CHECK("kb_jacobi_rc");
    return AJ_OK;
  }, st);
}

// per instance: sum the residual partials of layout w, record the norm and
// set the gate
int resnorm_final(aj_batch* b, const BatchWork& w, int per_item, hipStream_t st) {
  hipLaunchKernelGGL(kb_resnorm_final, dim3(b->ninst), dim3(64), 0, st,
                     (const double*)b->partial, w, per_item, b->res_stride,
                     (const double*)b->dxdy, (const double*)b->tol2, b->stats);
  AJ_LAUNCH_CHECK("kb_resnorm_final");
  return AJ_OK;
}

// One V-cycle from level lev down.  With eval_norm (level 0, p_src the
// iterate after the previous cycle), the pre-smoothing launch also sums the
// squared residual of p_src per tile, and the norm check after the previous
// cycle (its record and gate) runs right after it, before anything of this
// cycle writes p: an instance that converged with the previous cycle then
// skips the rest of this one, and its pre-smoothed scratch is not used.
template <typename T>
int batch_vcycle(aj_batch* b, int lev, const T* p_src, T* p_dst, const T* bb, int bzl,
                 hipStream_t st, bool eval_norm = false) {
  BatchMgArgs<T> m;
  m.stats = b->stats;
  m.bzl = bzl;
  m.res_partial = nullptr;
  m.res_stride = 0;
  if (lev > 0 && lev == b->sub_lev) {
    // this level and all coarser ones in one launch, one instance per
    // workgroup in LDS (bzl = 0 and a zero guess below level 0)
    BatchWork w = make_work(b, lev, 3);
    SubDims d;
    std::memset(&d, 0, sizeof(d));
    d.nl = b->ml - lev + 1;
    for (int g = 0; g < b->ng; ++g)
      for (int q = 0; q < d.nl; ++q) {
        d.nx[g][q] = b->lnx[g][lev + q];
        d.ny[g][q] = b->lny[g][lev + q];
      }
    m.p_in = nullptr;
    m.e_c = nullptr;
    m.b = bb;
    m.p_out = p_dst;
    m.clamp = 1;
    hipLaunchKernelGGL(kb_subcycle<T>, dim3(persistent_grid(b, w.total, kWgCoarse)),
                       dim3(kCoarseThreads), b->sub_lds, st, w, d, m);
    AJ_LAUNCH_CHECK("kb_subcycle");
    return AJ_OK;
  }
  if (lev >= b->ml) {
    // coarsest: 10 sweeps from zero; one workgroup per instance when it fits
    BatchWork w = make_work(b, lev, 3);
    long nmax = 0;
    for (int g = 0; g < b->ng; ++g) nmax = std::max<long>(nmax, (long)b->lnx[g][lev] * b->lny[g][lev]);
    if (nmax <= kCoarseThreads * kCoarsePerThread && nmax * 2 * (long)sizeof(T) <= 65536) {
      m.p_in = nullptr;
      m.e_c = nullptr;
      m.b = bb;
      m.p_out = p_dst;
      m.clamp = 0;
      hipLaunchKernelGGL(kb_coarse<T>, dim3(persistent_grid(b, w.total, kWgCoarse)), dim3(kCoarseThreads),
                         (size_t)(2 * nmax * sizeof(T)), st, w, m);
      AJ_LAUNCH_CHECK("kb_coarse");
      return AJ_OK;
    }
    // larger coarsest grids: 5 fused pairs of Jacobi sweeps, ping-pong
    T* other = (T*)b->Y[lev];
    const T* src = nullptr;
    for (int k = 0; k < 5; ++k) {
      T* dst = (k % 2 == 0) ? p_dst : other;
      m.p_in = src;
      m.e_c = nullptr;
      m.b = bb;
      m.p_out = dst;
      m.clamp = 0;
      if (int rc = batch_jacobi<T>(b, lev, "jacobi", m, st)) return rc;
      src = dst;
    }
    return AJ_OK;
  }
  T* Yl = (T*)b->Y[lev];
  m.p_in = p_src;
  m.e_c = nullptr;
  m.b = bb;
  m.p_out = Yl;
  m.clamp = 1;
  // algorithmic bytes of this level's launches (all instances)
  const double cl = (double)b->level_len[lev] * sizeof(T);
  if constexpr (std::is_same<T, float>::value) {
    if (jac_rc_ok(b, lev)) {
      // register columns: smoothing with the residual and its restriction
      // in one launch, then the coarser levels, then post-smoothing
      T* Bc = (T*)b->B[lev + 1];
      BatchWork wt;
      {
        // p (warm start) and b read, p and the quarter-size coarse rhs written
        ProfScope prof(lev == 0 ? AJ_PROF_MG_PRE0 : AJ_PROF_MG_PRE, cl * (p_src ? 3.25 : 2.25),
                       st);
        if (eval_norm && p_src) {
          m.res_partial = b->partial;
          m.res_stride = b->res_stride;
        }
        if (int rc = batch_jacobi_rc<true>(b, lev, "jacobi_rc", m, Bc, st, &wt)) return rc;
        m.res_partial = nullptr;
      }
      if (eval_norm && p_src)
        if (int rc = resnorm_final(b, wt, kJacRcW, st)) return rc;
      T* Ec = (T*)b->E[lev + 1];
      if (int rc = batch_vcycle<T>(b, lev + 1, nullptr, Ec, Bc, 0, st)) return rc;
      m.p_in = Yl;
      m.e_c = Ec;
      m.p_out = p_dst;
      m.clamp = 1;
      ProfScope prof(lev == 0 ? AJ_PROF_MG_POST0 : AJ_PROF_MG_POST, cl * 3.25, st);
      return batch_jacobi_rc<false>(b, lev, "jacobi_rc_post", m, nullptr, st);
    }
  }
  BatchWork wt;
  {
    // p (warm start) and b read, p written
    ProfScope prof(lev == 0 ? AJ_PROF_MG_PRE0 : AJ_PROF_MG_PRE, cl * (p_src ? 3 : 2), st);
    if (eval_norm && p_src) {
      m.res_partial = b->partial;
      m.res_stride = b->res_stride;
    }
    if (int rc = batch_jacobi<T>(b, lev, "jacobi", m, st, &wt)) return rc;
    m.res_partial = nullptr;
  }
  if (eval_norm && p_src)
    if (int rc = resnorm_final(b, wt, 4, st)) return rc;
  T* Bc = (T*)b->B[lev + 1];
  BatchWork wr = make_work(b, lev, 2);
  BatchMgArgs<T> r = m;
  r.p_out = Yl;
  {
    // residual + restriction: p, b read, the quarter-size coarse rhs written
    ProfScope prof(AJ_PROF_MG_RESID, cl * 2.25, st);
    if (int rc = batch_launch(b, "resrestrict", wr.total, [&](dim3 g, hipStream_t s) -> int {
          hipLaunchKernelGGL(kb_resrestrict<T>, g, dim3(256), 0, s, wr, r, Bc);
          AJ_LAUNCH_CHECK("kb_resrestrict");
          return AJ_OK;
        }, st))
      return rc;
  }
  T* Ec = (T*)b->E[lev + 1];
  if (int rc = batch_vcycle<T>(b, lev + 1, nullptr, Ec, Bc, 0, st)) return rc;
  m.p_in = Yl;
  m.e_c = Ec;
  m.p_out = p_dst;
  m.clamp = 1;
  // Y, b and the quarter-size coarse correction read, p written
  ProfScope prof(lev == 0 ? AJ_PROF_MG_POST0 : AJ_PROF_MG_POST, cl * 3.25, st);
  return batch_jacobi<T>(b, lev, "jacobi_post", m, st);
}

template <typename T>
BatchStepArgs<T> step_args(aj_batch* b, const aj_batch_step_params* prm, const void* u,
                           const void* v, const void* mask, const void* sdf, void* rhs, void* p,
                           void* u_out, void* v_out) {
  BatchStepArgs<T> s;
  s.u = (const T*)u;
  s.v = (const T*)v;
  s.mask = (const T*)mask;
  s.sdf = (const T*)sdf;
  s.us = (T*)b->us;
  s.vs = (T*)b->vs;
  s.rhs = (T*)rhs;
  s.p = (T*)p;
  s.uo = (T*)u_out;
  s.vo = (T*)v_out;
  s.nsgs = prm->smagorinsky_cs > 0 ? (T*)b->nsgs : nullptr;
  s.nu = b->nu;
  s.dt = prm->dt;
  s.U_inf = prm->U_inf;
  s.inv_eta = 1.0 / prm->brinkman_eta;
  s.cs = prm->smagorinsky_cs;
  return s;
}

// one register-column RK launch of W waves x R rows per wave: tiles of U
// rows x (the whole row, or UY-column strips past 128 columns)
template <int W, int R>
int batch_rk_rc(aj_batch* b, const BatchStepArgs<float>& s, hipStream_t st) {
  using G = RkRcGeo<W, R, false, true>;
  BatchWork w = make_work(b, 0, 0, G::U, 128);
  {
    int tot = 0;
    for (int g = 0; g < w.ng; ++g) {
      const int ny = w.ny[g];
      w.tiles_y[g] = ny <= 128 ? 1 : (int)cdiv(ny, G::UY);
      w.per_inst[g] = G::x_tiles(w.nx[g]) * w.tiles_y[g];
      w.wstart[g] = tot;
      tot += w.per_inst[g] * b->gcount[g];
    }
    w.wstart[w.ng] = tot;
    w.total = tot;
  }
  return batch_launch(b, "rk_rc", w.total, [&](dim3 g, hipStream_t t) -> int {
    hipLaunchKernelGGL((kb_rk_rc<W, R>), g, dim3(W * 64), 0, t, w, s);
    AJ_LAUNCH_CHECK("kb_rk_rc");
    return AJ_OK;
  }, st);
}

// [LES] + RK3 predictor (u*, v* in the batch's buffers) + rhs
template <typename T>
int predict_launches(aj_batch* b, const BatchStepArgs<T>& s, hipStream_t st) {
  BatchWork wrow = make_work(b, 0, 1);
  if (s.nsgs) {
    ProfScope prof(AJ_PROF_LES, (double)b->cells * sizeof(T) * 3, st);
    if (int rc = batch_launch(b, "les", wrow.total, [&](dim3 g, hipStream_t t) -> int {
          hipLaunchKernelGGL(kb_les<T>, g, dim3(256), 0, t, wrow, s);
          AJ_LAUNCH_CHECK("kb_les");
          return AJ_OK;
        }, st))
      return rc;
  }
  if constexpr (std::is_same<T, float>::value) {
    bool rc_ok = true;
    for (int g = 0; g < b->ng; ++g) {
      const int nx = b->lnx[g][0], ny = b->lny[g][0];
      if ((ny & 1) || ny < 4 || nx < 4) rc_ok = false;
    }
    if (rc_ok) {
      // register columns: RK3 predictor and the pressure rhs in one launch
      ProfScope prof(AJ_PROF_RK, (double)b->cells * sizeof(T) * (s.nsgs ? 8 : 7), st);
      return batch_rk_rc<kRkRcW, kRkRcR>(b, s, st);
    }
  }
  // RK3 predictor (the predictor BCs of step.py:182-191 are its own last step)
  {
    constexpr int TX = BatchRkTile<T>::TX, TY = BatchRkTile<T>::TY, H = 3;
    BatchWork w = make_work(b, 0, 0, TX, TY);
    const size_t smem = (size_t)(TX + 2 * H) * (TY + 2 * H) * 8 * sizeof(T);
    static_assert((TX + 2 * H) * (TY + 2 * H) * 8 * sizeof(T) <= 65536, "RK tile LDS");
    ProfScope prof(AJ_PROF_RK, (double)b->cells * sizeof(T) * (s.nsgs ? 7 : 6), st);
    if (int rc = batch_launch(b, "rk", w.total, [&](dim3 g, hipStream_t t) -> int {
          hipLaunchKernelGGL((kb_rk<T, TX, TY, kBatchRkNTH>), g, dim3(kBatchRkNTH), smem, t, w, s);
          AJ_LAUNCH_CHECK("kb_rk");
          return AJ_OK;
        }, st))
      return rc;
  }
  ProfScope prof(AJ_PROF_RHS, (double)b->cells * sizeof(T) * 4, st);
  return batch_launch(b, "div", wrow.total, [&](dim3 g, hipStream_t t) -> int {
    hipLaunchKernelGGL(kb_div<T>, g, dim3(256), 0, t, wrow, s);
    AJ_LAUNCH_CHECK("kb_div");
    return AJ_OK;
  }, st);
}

// reference multigrid, v_cycles cycles with the per-instance gate
template <typename T>
int refmg_launches(aj_batch* b, const aj_batch_step_params* prm, const T* rhs, T* p_out,
                   hipStream_t st) {
  hipLaunchKernelGGL(kb_mg_begin, dim3(b->ninst), dim3(64), 0, st, b->stats, b->ninst);
  AJ_LAUNCH_CHECK("kb_mg_begin");
  // the norm after cycle c < v_cycles - 1 is evaluated inside cycle c + 1
  // (batch_vcycle's eval_norm); the one after the last cycle by its own pass
  for (int c = 0; c < prm->v_cycles; ++c) {
    const T* src = (c == 0) ? nullptr : p_out;
    if (int rc = batch_vcycle<T>(b, 0, src, p_out, rhs, 1, st, c > 0)) return rc;
  }
  if (prm->v_cycles <= 0) return AJ_OK;
  BatchWork wn = make_work(b, 0, 4);
  if (int rc = batch_launch(b, "resnorm", wn.total, [&](dim3 g, hipStream_t t) -> int {
        hipLaunchKernelGGL(kb_resnorm<T>, g, dim3(256), 0, t, wn, (const T*)p_out, rhs,
                           b->partial, b->res_stride);
        AJ_LAUNCH_CHECK("kb_resnorm");
        return AJ_OK;
      }, st))
    return rc;
  return resnorm_final(b, wn, 1, st);
}

template <typename T>
int project_launches(aj_batch* b, const BatchStepArgs<T>& s, hipStream_t st) {
  BatchWork wrow = make_work(b, 0, 1);
  ProfScope prof(AJ_PROF_PROJECT, (double)b->cells * sizeof(T) * 6, st);
  int pairs = 1;
  for (const void* a : {(const void*)s.us, (const void*)s.vs, (const void*)s.p,
                        (const void*)s.mask, (const void*)s.uo, (const void*)s.vo})
    if ((uintptr_t)a % (2 * sizeof(T))) pairs = 0;
  return batch_launch(b, "project", wrow.total, [&](dim3 g, hipStream_t t) -> int {
    hipLaunchKernelGGL(kb_project<T>, g, dim3(256), 0, t, wrow, s, pairs);
    AJ_LAUNCH_CHECK("kb_project");
    return AJ_OK;
  }, st);
}

template <typename T>
int batch_step_launches(aj_batch* b, const aj_batch_step_params* prm, const void* u,
                        const void* v, const void* mask, const void* sdf, void* u_out,
                        void* v_out, void* p_out, hipStream_t st) {
  const BatchStepArgs<T> s = step_args<T>(b, prm, u, v, mask, sdf, b->rhs, p_out, u_out, v_out);
  if (int rc = predict_launches<T>(b, s, st)) return rc;
  if (int rc = refmg_launches<T>(b, prm, (const T*)b->rhs, (T*)p_out, st)) return rc;
  return project_launches<T>(b, s, st);
}

// tolerance**2 in the working precision (multigrid_solver.py:229-231)
int set_tolerance(aj_batch* b, double tol, hipStream_t st) {
  if (tol == b->last_tol) return AJ_OK;
  std::vector<double> t2(b->ninst);
  const double tt = b->dtype == AJ_F64 ? tol * tol : (double)((float)tol * (float)tol);
  for (auto& x : t2) x = tt;
  AJ_HIP(hipMemcpyAsync(b->tol2, t2.data(), t2.size() * 8, hipMemcpyHostToDevice, st));
  AJ_HIP(hipStreamSynchronize(st));
  b->last_tol = tol;
  return AJ_OK;
}

// textbook solve of every group to prm->tolerance; p0 may be null
int textbook_solve(aj_batch* b, const void* rhs, const void* p0, void* p_out,
                   const aj_batch_step_params* prm, hipStream_t st) {
  const size_t es = b->dtype == AJ_F64 ? 8 : 4;
  if (b->tb.empty()) {
    b->tb.assign(b->ng, nullptr);
    for (int g = 0; g < b->ng; ++g)
      if (int rc = aj_mg_create_batched(&b->tb[g], b->dtype, b->gnx[g], b->gny[g], b->gdx[g],
                                        b->gdy[g], 0, b->gcount[g]))
        return rc;
  }
  std::vector<const void*> r(b->ng), g0(b->ng);
  std::vector<void*> o(b->ng);
  for (int g = 0; g < b->ng; ++g) {
    const size_t off = (size_t)b->lbase[0][g] * es;
    r[g] = (const char*)rhs + off;
    g0[g] = p0 ? (const char*)p0 + off : nullptr;
    o[g] = (char*)p_out + off;
  }
  b->last_solver = AJ_MG_TEXTBOOK;
  return tb_solve_many(b->tb.data(), b->ng, r.data(), g0.data(), o.data(), prm->v_cycles,
                       prm->tolerance, st);
}

int check_params(const aj_batch_step_params* prm, const char* who) {
  char msg[160];
  if (!(prm->brinkman_eta > 0) || prm->v_cycles < 1) {
    std::snprintf(msg, sizeof(msg), "%s: brinkman_eta > 0 and v_cycles >= 1 required", who);
    return fail(AJ_ERR_INVALID, msg);
  }
  if (prm->solver != AJ_MG_REFERENCE && prm->solver != AJ_MG_TEXTBOOK) {
    std::snprintf(msg, sizeof(msg), "%s: solver must be AJ_MG_REFERENCE or AJ_MG_TEXTBOOK", who);
    return fail(AJ_ERR_INVALID, msg);
  }
  if (!(prm->smagorinsky_cs >= 0)) {
    std::snprintf(msg, sizeof(msg), "%s: smagorinsky_cs must be >= 0", who);
    return fail(AJ_ERR_INVALID, msg);
  }
  return AJ_OK;
}

}  // namespace

extern "C" int aj_batch_create(aj_batch** out, int dtype, int ngroups, const int* nx,
                               const int* ny, const int* count, const double* dx,
                               const double* dy, int levels) {
  if (!out || !nx || !ny || !count || !dx || !dy) return fail(AJ_ERR_INVALID, "aj_batch_create: null pointer");
  *out = nullptr;
  if (dtype != AJ_F32 && dtype != AJ_F64) return fail(AJ_ERR_INVALID, "aj_batch_create: bad dtype");
  if (ngroups < 1 || ngroups > kMaxGroups)
    return fail(AJ_ERR_INVALID, "aj_batch_create: 1..8 groups of equal grids");
  auto* b = new aj_batch();
  b->dtype = dtype;
  b->ng = ngroups;
  b->levels = levels;
  int ml_all = -1;
  for (int g = 0; g < ngroups; ++g) {
    if (nx[g] < 4 || ny[g] < 4 || count[g] < 1 || !(dx[g] > 0) || !(dy[g] > 0)) {
      delete b;
      return fail(AJ_ERR_INVALID, "aj_batch_create: bad group");
    }
    // reference hierarchy (multigrid_solver.py:96-106)
    int ml = 1, tx = nx[g], ty = ny[g];
    while (tx % 2 == 0 && ty % 2 == 0 && ml < levels) {
      tx /= 2;
      ty /= 2;
      ++ml;
    }
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "Grid dimensions (%d, %d) not suitable for multigrid with %d levels. Grid must "
                  "be divisible by 2^%d",
                  nx[g], ny[g], levels, levels);
    if (ml < 2) {
      delete b;
      return fail(AJ_ERR_GRID, msg);
    }
    if (ml_all >= 0 && ml != ml_all) {
      delete b;
      return fail(AJ_ERR_INVALID, "aj_batch_create: all groups need the same multigrid depth");
    }
    ml_all = ml;
    std::vector<int> sx{nx[g]}, sy{ny[g]};
    for (int l = 0; l < ml; ++l) {
      int cx = sx.back() / 2, cy = sy.back() / 2;
      if (cx < 2 || cy < 2) cx = cy = 1;
      sx.push_back(cx);
      sy.push_back(cy);
    }
    b->lnx.push_back(sx);
    b->lny.push_back(sy);
    b->gnx.push_back(nx[g]);
    b->gny.push_back(ny[g]);
    b->gcount.push_back(count[g]);
    b->gdx.push_back(dx[g]);
    b->gdy.push_back(dy[g]);
    b->gfirst.push_back(b->ninst);
    b->ninst += count[g];
  }
  b->ml = ml_all;
  const int L = b->ml;
  b->lbase.assign(L + 1, std::vector<long>(ngroups, 0));
  b->level_len.assign(L + 1, 0);
  for (int l = 0; l <= L; ++l) {
    long off = 0;
    for (int g = 0; g < ngroups; ++g) {
      b->lbase[l][g] = off;
      off += (long)b->lnx[g][l] * b->lny[g][l] * b->gcount[g];
    }
    b->level_len[l] = off;
  }
  b->cells = b->level_len[0];
  const size_t es = dtype == AJ_F64 ? 8 : 4;
  size_t bytes = 0;
  auto take = [&](size_t n) {
    const size_t o = bytes;
    bytes += (n + 255) & ~size_t(255);
    return o;
  };
  const size_t o_us = take(b->cells * es), o_vs = take(b->cells * es), o_rhs = take(b->cells * es);
  const size_t o_nsgs = take(b->cells * es);
  std::vector<size_t> oB(L + 1), oY(L + 1), oE(L + 1);
  for (int l = 0; l <= L; ++l) {
    oB[l] = take(b->level_len[l] * es);
    oY[l] = take(b->level_len[l] * es);
    oE[l] = take(b->level_len[l] * es);
  }
  // partial sums per instance: kResChunks (norm pass) or four per level-0
  // Jacobi tile (any of the three tile shapes)
  b->res_stride = kResChunks;
  for (int g = 0; g < ngroups; ++g) {
    const long nx = b->lnx[g][0], ny = b->lny[g][0];
    const long t = std::max({cdiv(nx, kBatchJX) * cdiv(ny, kBatchJY), cdiv(nx, 32) * cdiv(ny, 32),
                             cdiv(nx, 64) * cdiv(ny, 16)});
    b->res_stride = (int)std::max<long>(b->res_stride, 4 * t);
    // register-column tiles: one partial per wave
    const long trc = cdiv(nx, JacRcGeo<kJacRcW, kJacRcR, true>::U);
    b->res_stride = (int)std::max<long>(b->res_stride, kJacRcW * trc);
  }
  const size_t o_part = take((size_t)b->ninst * b->res_stride * 8);
  const size_t o_dxdy = take((size_t)b->ninst * 8), o_tol2 = take((size_t)b->ninst * 8);
  const size_t o_nu = take((size_t)b->ninst * 8);
  const size_t o_obs = take((size_t)b->ninst * sizeof(ObstacleDev));
  const size_t o_st = take((size_t)b->ninst * sizeof(aj_mg_stats));
  hipError_t e = hipGetDevice(&b->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device);
  // the in-LDS V-cycle starts at the first level below 0 whose levels fit
  // one workgroup's LDS (w0, w1 of the top level, b and y of every level)
  for (int l = 1; l < L && b->ml - l + 1 <= kMaxRefLevels; ++l) {
    size_t need = 0;
    for (int g = 0; g < ngroups; ++g) {
      const size_t n0 = (size_t)b->lnx[g][l] * b->lny[g][l];
      size_t tot = 2 * n0;
      for (int q = l; q <= L; ++q) tot += 2 * (size_t)b->lnx[g][q] * b->lny[g][q];
      need = std::max(need, tot * es);
    }
    if (need <= 160 * 1024) {
      b->sub_lev = l;
      b->sub_lds = need;
      break;
    }
  }
  if (e == hipSuccess && b->sub_lds > 65536)
    e = hipFuncSetAttribute(dtype == AJ_F64 ? (const void*)kb_subcycle<double>
                                            : (const void*)kb_subcycle<float>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)b->sub_lds);
  if (e == hipSuccess) e = hipMalloc(&b->ws, bytes);
  if (e != hipSuccess) {
    delete b;
    return hip_fail(e, "aj_batch_create");
  }
  char* base = (char*)b->ws;
  b->us = base + o_us;
  b->vs = base + o_vs;
  b->rhs = base + o_rhs;
  b->nsgs = base + o_nsgs;
  b->B.assign(L + 1, nullptr);
  b->Y.assign(L + 1, nullptr);
  b->E.assign(L + 1, nullptr);
  for (int l = 0; l <= L; ++l) {
    b->B[l] = base + oB[l];
    b->Y[l] = base + oY[l];
    b->E[l] = base + oE[l];
  }
  b->partial = (double*)(base + o_part);
  b->dxdy = (double*)(base + o_dxdy);
  b->tol2 = (double*)(base + o_tol2);
  b->nu = (double*)(base + o_nu);
  b->obs = base + o_obs;
  b->stats = (aj_mg_stats*)(base + o_st);
  std::vector<double> dd(b->ninst);
  for (int g = 0; g < ngroups; ++g)
    for (int k = 0; k < b->gcount[g]; ++k) {
      const double ddx = dtype == AJ_F64 ? dx[g] : (double)(float)dx[g];
      const double ddy = dtype == AJ_F64 ? dy[g] : (double)(float)dy[g];
      dd[b->gfirst[g] + k] = ddx * ddy;
    }
  e = hipMemcpy(b->dxdy, dd.data(), dd.size() * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(b->stats, 0, (size_t)b->ninst * sizeof(aj_mg_stats));
  if (e != hipSuccess) {
    (void)hipFree(b->ws);
    delete b;
    return hip_fail(e, "aj_batch_create");
  }
  *out = b;
  return AJ_OK;
}

extern "C" int aj_batch_destroy(aj_batch* b) {
  if (!b) return AJ_OK;
  for (int k = 0; k < 2; ++k) {
    if (b->exec[k]) (void)hipGraphExecDestroy(b->exec[k]);
    if (b->graph[k]) (void)hipGraphDestroy(b->graph[k]);
  }
  for (aj_mg_plan* p : b->tb)
    if (p) (void)aj_mg_destroy(p);
  if (b->ws) (void)hipFree(b->ws);
  delete b;
  return AJ_OK;
}

extern "C" long aj_batch_cells(const aj_batch* b) { return b ? b->cells : 0; }
extern "C" int aj_batch_num_instances(const aj_batch* b) { return b ? b->ninst : 0; }

extern "C" int aj_batch_offsets(const aj_batch* b, long* offsets) {
  if (!b || !offsets) return fail(AJ_ERR_INVALID, "aj_batch_offsets: null pointer");
  for (int g = 0; g < b->ng; ++g)
    for (int k = 0; k < b->gcount[g]; ++k)
      offsets[b->gfirst[g] + k] = b->lbase[0][g] + (long)k * b->gnx[g] * b->gny[g];
  return AJ_OK;
}

extern "C" int aj_batch_set_viscosity(aj_batch* b, const double* nu_host) {
  if (!b || !nu_host) return fail(AJ_ERR_INVALID, "aj_batch_set_viscosity: null pointer");
  AJ_HIP(hipMemcpy(b->nu, nu_host, (size_t)b->ninst * 8, hipMemcpyHostToDevice));
  return AJ_OK;
}

extern "C" int aj_batch_set_option(aj_batch* b, int option, int value) {
  if (!b) return fail(AJ_ERR_INVALID, "aj_batch_set_option: null batch");
  if (option != AJ_MG_OPT_GRAPH) return fail(AJ_ERR_INVALID, "aj_batch_set_option: unknown option");
  b->use_graph = value ? 1 : 0;
  return AJ_OK;
}

extern "C" int aj_batch_step(aj_batch* b, const void* u, const void* v, const void* mask,
                             const void* sdf, void* u_out, void* v_out, void* p_out,
                             const aj_batch_step_params* prm, void* stream) {
  if (!b || !u || !v || !mask || !u_out || !v_out || !p_out || !prm)
    return fail(AJ_ERR_INVALID, "aj_batch_step: null pointer");
  if (u_out == u || v_out == v) return fail(AJ_ERR_INVALID, "aj_batch_step: outputs must not alias inputs");
  if (int rc = check_params(prm, "aj_batch_step")) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (prm->solver == AJ_MG_TEXTBOOK) {
    // predictor and projection around the to-tolerance solve (its outer
    // iterations replay the plans' captured graphs)
    if (int rc = aj_batch_predict(b, u, v, mask, sdf, b->rhs, prm, stream)) return rc;
    if (int rc = aj_batch_solve(b, b->rhs, nullptr, p_out, prm, stream)) return rc;
    return aj_batch_project(b, p_out, mask, u_out, v_out, prm, stream);
  }
  if (int rc = set_tolerance(b, prm->tolerance, st)) return rc;
  b->last_solver = AJ_MG_REFERENCE;
  auto launches = [&](hipStream_t s) {
    return b->dtype == AJ_F32
               ? batch_step_launches<float>(b, prm, u, v, mask, sdf, u_out, v_out, p_out, s)
               : batch_step_launches<double>(b, prm, u, v, mask, sdf, u_out, v_out, p_out, s);
  };
  // the first step runs eagerly so that every launch shape is tuned before
  // the step is captured
  if (!b->use_graph || prof_on() || !b->tuned) {
    const int rc = launches(st);
    b->tuned = 1;
    return rc;
  }
  const void* key[8] = {u, v, mask, sdf, u_out, v_out, p_out, nullptr};
  const double kp[7] = {prm->dt,        prm->U_inf,         prm->brinkman_eta, (double)prm->v_cycles,
                        prm->tolerance, prm->smagorinsky_cs, (double)prm->solver};
  int slot = -1;
  for (int k = 0; k < 2; ++k)
    if (b->exec[k] && !std::memcmp(key, b->key[k], sizeof(key)) &&
        !std::memcmp(kp, b->key_prm[k], sizeof(kp)))
      slot = k;
  if (slot < 0) {
    slot = b->victim;
    b->victim ^= 1;
    if (b->exec[slot]) {
      (void)hipGraphExecDestroy(b->exec[slot]);
      (void)hipGraphDestroy(b->graph[slot]);
      b->exec[slot] = nullptr;
      b->graph[slot] = nullptr;
    }
    hipStream_t cs;
    AJ_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    int rc = AJ_OK;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      rc = launches(cs);
      hipGraph_t g = nullptr;
      e = hipStreamEndCapture(cs, &g);
      if (e == hipSuccess && rc == AJ_OK) {
        e = hipGraphInstantiate(&b->exec[slot], g, nullptr, nullptr, 0);
        b->graph[slot] = g;
      } else if (g) {
        (void)hipGraphDestroy(g);
      }
    }
    (void)hipStreamDestroy(cs);
    if (rc != AJ_OK) return rc;
    if (e != hipSuccess) return hip_fail(e, "capture of the batched step");
    std::memcpy(b->key[slot], key, sizeof(key));
    std::memcpy(b->key_prm[slot], kp, sizeof(kp));
  }
  AJ_HIP(hipGraphLaunch(b->exec[slot], st));
  return AJ_OK;
}

extern "C" int aj_batch_predict(aj_batch* b, const void* u, const void* v, const void* mask,
                                const void* sdf, void* rhs, const aj_batch_step_params* prm,
                                void* stream) {
  if (!b || !u || !v || !mask || !rhs || !prm)
    return fail(AJ_ERR_INVALID, "aj_batch_predict: null pointer");
  if (int rc = check_params(prm, "aj_batch_predict")) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == AJ_F32)
    return predict_launches<float>(
        b, step_args<float>(b, prm, u, v, mask, sdf, rhs, nullptr, nullptr, nullptr), st);
  return predict_launches<double>(
      b, step_args<double>(b, prm, u, v, mask, sdf, rhs, nullptr, nullptr, nullptr), st);
}

extern "C" int aj_batch_solve(aj_batch* b, const void* rhs, const void* p0, void* p_out,
                              const aj_batch_step_params* prm, void* stream) {
  if (!b || !rhs || !p_out || !prm) return fail(AJ_ERR_INVALID, "aj_batch_solve: null pointer");
  if (rhs == p_out || (p0 && p0 == p_out))
    return fail(AJ_ERR_INVALID, "aj_batch_solve: p_out must not alias rhs or p0");
  if (int rc = check_params(prm, "aj_batch_solve")) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (prm->solver == AJ_MG_TEXTBOOK) return textbook_solve(b, rhs, p0, p_out, prm, st);
  if (p0) return fail(AJ_ERR_INVALID, "aj_batch_solve: the reference solver starts from zero");
  if (int rc = set_tolerance(b, prm->tolerance, st)) return rc;
  b->last_solver = AJ_MG_REFERENCE;
  if (b->dtype == AJ_F32) return refmg_launches<float>(b, prm, (const float*)rhs, (float*)p_out, st);
  return refmg_launches<double>(b, prm, (const double*)rhs, (double*)p_out, st);
}

extern "C" int aj_batch_project(aj_batch* b, const void* p, const void* mask, void* u_out,
                                void* v_out, const aj_batch_step_params* prm, void* stream) {
  if (!b || !p || !mask || !u_out || !v_out || !prm)
    return fail(AJ_ERR_INVALID, "aj_batch_project: null pointer");
  if (int rc = check_params(prm, "aj_batch_project")) return rc;
  hipStream_t st = (hipStream_t)stream;
  if (b->dtype == AJ_F32)
    return project_launches<float>(
        b, step_args<float>(b, prm, nullptr, nullptr, mask, nullptr, nullptr, (void*)p, u_out, v_out),
        st);
  return project_launches<double>(
      b, step_args<double>(b, prm, nullptr, nullptr, mask, nullptr, nullptr, (void*)p, u_out, v_out),
      st);
}

namespace aj {
void obstacle_rotation(const aj_obstacle& o, double* ca, double* sa);
int check_obstacle(const aj_obstacle& o, const char* who);
}  // namespace aj

extern "C" int aj_batch_geometry(aj_batch* b, const aj_obstacle* obstacles, void* sdf, void* mask,
                                 void* stream) {
  if (!b || !obstacles || (!sdf && !mask)) return fail(AJ_ERR_INVALID, "aj_batch_geometry: null pointer");
  std::vector<ObstacleDev> h(b->ninst);
  for (int k = 0; k < b->ninst; ++k) {
    if (int rc = check_obstacle(obstacles[k], "aj_batch_geometry")) return rc;
    h[k].o = obstacles[k];
    obstacle_rotation(obstacles[k], &h[k].ca, &h[k].sa);
  }
  hipStream_t st = (hipStream_t)stream;
  AJ_HIP(hipMemcpyAsync(b->obs, h.data(), h.size() * sizeof(ObstacleDev), hipMemcpyHostToDevice, st));
  BatchWork w = make_work(b, 0, 1);
  const dim3 grid(persistent_grid(b, w.total, 8));
  if (b->dtype == AJ_F32)
    hipLaunchKernelGGL(kb_geometry<float>, grid, dim3(256), 0, st, w, (const ObstacleDev*)b->obs,
                       (float*)sdf, (float*)mask);
  else
    hipLaunchKernelGGL(kb_geometry<double>, grid, dim3(256), 0, st, w, (const ObstacleDev*)b->obs,
                       (double*)sdf, (double*)mask);
  AJ_LAUNCH_CHECK("kb_geometry");
  // the host array must outlive the copy
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}

extern "C" int aj_batch_get_stats(aj_batch* b, aj_mg_stats* host, void* stream) {
  if (!b || !host) return fail(AJ_ERR_INVALID, "aj_batch_get_stats: null pointer");
  hipStream_t st = (hipStream_t)stream;
  if (b->last_solver == AJ_MG_TEXTBOOK && !b->tb.empty()) {
    // instances are numbered group by group, as the plans hold them
    for (int g = 0; g < b->ng; ++g)
      AJ_HIP(hipMemcpyAsync(host + b->gfirst[g], b->tb[g]->stats,
                            (size_t)b->gcount[g] * sizeof(aj_mg_stats), hipMemcpyDeviceToHost, st));
  } else {
    AJ_HIP(hipMemcpyAsync(host, b->stats, (size_t)b->ninst * sizeof(aj_mg_stats),
                          hipMemcpyDeviceToHost, st));
  }
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}
