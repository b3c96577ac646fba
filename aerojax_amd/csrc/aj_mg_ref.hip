// Reference-exact geometric multigrid (pressure_solvers/multigrid_solver.py:55-241)
// and the multigrid plan API.
//
// One V-cycle at level l < max_levels is two kernels:
//   pre : 2 Jacobi sweeps (temporal blocking in LDS, halo 2) + outlet clamp
//   rr  : residual b - A p and 2x2 box restriction, one thread per coarse cell
// then, after the coarser levels,
//   post: load p + prolong(e_coarse) on the tile + halo, 2 Jacobi sweeps,
//         outlet clamp.
// The coarsest level (10 sweeps from zero) runs in one workgroup in LDS.
// After each cycle a fixed-order reduction evaluates sum(r^2)*dx*dy on the
// device and sets the `converged` gate that every later kernel of the solve
// reads first, so the whole solve is enqueued without a host round trip.
#include "aj_mg.h"

#include <cmath>
#include <cstring>

namespace aj {

constexpr int kNPartial = 1024;

template <typename T>
struct RefJacobiArgs {
  const T* p_in;  // null -> zero initial guess
  const T* e_c;   // null -> no prolongation; else coarse correction
  int ncx, ncy;
  const T* b;
  int b_zero_last;  // level 0: b[-1,:] = 0
  T* p_out;
  int nx, ny;
  T ax, ay, inv_diag;
  int clamp_last;
  const aj_mg_stats* gate;
};

__device__ __forceinline__ bool gated(const aj_mg_stats* g) {
  return g != nullptr && *(volatile const int*)&g->converged != 0;
}

// prolong (multigrid_solver.py:134-156) evaluated at fine cell (i, j)
template <typename T>
__device__ __forceinline__ T prolong_ref(const T* __restrict__ c, int ncx, int ncy, int i, int j) {
  const int nxf = 2 * ncx, nyf = 2 * ncy;
  if (i == nxf - 1 || j == nyf - 1) return T(0);
  const int I = i >> 1, J = j >> 1;
  const bool io = i & 1, jo = j & 1;
  const T c00 = c[(long)I * ncy + J];
  if (!io && !jo) return c00;
  if (io && !jo) return T(0.5) * (c00 + c[(long)(I + 1) * ncy + J]);
  if (!io && jo) return T(0.5) * (c00 + c[(long)I * ncy + J + 1]);
  return T(0.25) * (((c00 + c[(long)(I + 1) * ncy + J]) + c[(long)I * ncy + J + 1]) +
                    c[(long)(I + 1) * ncy + J + 1]);
}

// NS Jacobi sweeps (smooth, multigrid_solver.py:158-171) on one tile.
template <typename T, int TX, int TY, int NS>
__global__ __launch_bounds__(256) void k_ref_jacobi(RefJacobiArgs<T> a) {
  constexpr int RH = TX + 2 * NS, RW = TY + 2 * NS, RN = RH * RW;
  __shared__ T s_p[2][RN];
  __shared__ T s_b[RN];
  const int nx = a.nx, ny = a.ny;
  const int gx0 = blockIdx.y * TX - NS, gy0 = blockIdx.x * TY - NS;
  const int tid = threadIdx.x;
  {
    // all loads of a thread in flight together (clamped, always in bounds)
    constexpr int NIT = (RN + 255) / 256;
    T rp[NIT], rb[NIT];
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int idx = min(tid + it * 256, RN - 1);
      const int lx = idx / RW, ly = idx - lx * RW;
      const int gx = clampi(gx0 + lx, 0, nx - 1), gy = clampi(gy0 + ly, 0, ny - 1);
      const long g = (long)gx * ny + gy;
      T pv = a.p_in ? a.p_in[g] : T(0);
      if (a.e_c) pv = pv + prolong_ref(a.e_c, a.ncx, a.ncy, gx, gy);
      rp[it] = pv;
      rb[it] = (a.b_zero_last && gx == nx - 1) ? T(0) : a.b[g];
    }
    if (gated(a.gate)) return;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int idx = tid + it * 256;
      if (idx < RN) {
        s_p[0][idx] = rp[it];
        s_b[idx] = rb[it];
      }
    }
  }
  __syncthreads();
  int cur = 0;
#pragma unroll
  for (int s = 1; s <= NS; ++s) {
    const int m = NS - s;
    const int W = TY + 2 * m, Hh = TX + 2 * m;
    const T* src = s_p[cur];
    T* dst = s_p[cur ^ 1];
    for (int idx = tid; idx < W * Hh; idx += 256) {
      const int lx = NS - m + idx / W, ly = NS - m + idx % W;
      const int gx = gx0 + lx, gy = gy0 + ly;
      if (gx < 0 || gx >= nx || gy < 0 || gy >= ny) continue;
      const int l = lx * RW + ly;
      const T pW = src[gx > 0 ? l - RW : l], pE = src[gx < nx - 1 ? l + RW : l];
      const T pS = src[gy > 0 ? l - 1 : l], pN = src[gy < ny - 1 ? l + 1 : l];
      dst[l] = (a.ax * (pE + pW) + a.ay * (pN + pS) - s_b[l]) * a.inv_diag;
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int idx = tid; idx < TX * TY; idx += 256) {
    const int lx = NS + idx / TY, ly = NS + idx % TY;
    const int gx = gx0 + lx, gy = gy0 + ly;
    if (gx >= nx || gy >= ny) continue;
    const T val = (a.clamp_last && gx == nx - 1) ? T(0) : s_p[cur][lx * RW + ly];
    a.p_out[(long)gx * ny + gy] = val;
  }
}

// b - apply_laplacian(p) at (i, j) (multigrid_solver.py:173-183, 198)
template <typename T>
__device__ __forceinline__ T ref_residual(const T* __restrict__ p, const T* __restrict__ b,
                                          int bzl, int nx, int ny, T ax, T ay, int i, int j) {
  const long g = (long)i * ny + j;
  const T pc = p[g];
  const T pW = i > 0 ? p[g - ny] : pc, pE = i < nx - 1 ? p[g + ny] : pc;
  const T pS = j > 0 ? p[g - 1] : pc, pN = j < ny - 1 ? p[g + 1] : pc;
  const T lap = ax * (pE + pW - T(2) * pc) + ay * (pN + pS - T(2) * pc);
  const T bv = (bzl && i == nx - 1) ? T(0) : b[g];
  return bv - lap;
}

// residual + restrict (multigrid_solver.py:108-132, 198-199)
template <typename T>
__global__ __launch_bounds__(256) void k_ref_resrestrict(const T* __restrict__ p,
                                                         const T* __restrict__ b, int bzl, int nx,
                                                         int ny, T ax, T ay, T* __restrict__ bc,
                                                         int ncx, int ncy, int degenerate,
                                                         const aj_mg_stats* gate) {
  if (gated(gate)) return;
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = blockIdx.y;
  if (J >= ncy || I >= ncx) return;
  if (degenerate) {  // coarse side < 2: fine[:1, :1]
    bc[0] = ref_residual(p, b, bzl, nx, ny, ax, ay, 0, 0);
    return;
  }
  const int i = 2 * I, j = 2 * J;
  const T r00 = ref_residual(p, b, bzl, nx, ny, ax, ay, i, j);
  const T r10 = ref_residual(p, b, bzl, nx, ny, ax, ay, i + 1, j);
  const T r01 = ref_residual(p, b, bzl, nx, ny, ax, ay, i, j + 1);
  const T r11 = ref_residual(p, b, bzl, nx, ny, ax, ay, i + 1, j + 1);
  bc[(long)I * ncy + J] = T(0.25) * (((r00 + r10) + r01) + r11);
}

// coarsest level: NSW Jacobi sweeps from zero in one workgroup
constexpr int kCoarseThreads = 1024;
constexpr int kCoarsePerThread = 8;

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_ref_coarse(const T* __restrict__ b,
                                                               T* __restrict__ out, int nx, int ny,
                                                               T ax, T ay, T inv_diag, int nsw,
                                                               const aj_mg_stats* gate) {
  if (gated(gate)) return;
  extern __shared__ unsigned char smem_raw[];
  T* s0 = reinterpret_cast<T*>(smem_raw);
  const int n = nx * ny;
  T* s1 = s0 + n;
  T breg[kCoarsePerThread];
#pragma unroll
  for (int k = 0; k < kCoarsePerThread; ++k) {
    const int c = threadIdx.x + k * kCoarseThreads;
    breg[k] = c < n ? b[c] : T(0);
    if (c < n) s0[c] = T(0);
  }
  __syncthreads();
  T* src = s0;
  T* dst = s1;
  for (int s = 0; s < nsw; ++s) {
#pragma unroll
    for (int k = 0; k < kCoarsePerThread; ++k) {
      const int c = threadIdx.x + k * kCoarseThreads;
      if (c < n) {
        const int i = c / ny, j = c - i * ny;
        const T pc = src[c];
        const T pW = i > 0 ? src[c - ny] : pc, pE = i < nx - 1 ? src[c + ny] : pc;
        const T pS = j > 0 ? src[c - 1] : pc, pN = j < ny - 1 ? src[c + 1] : pc;
        dst[c] = (ax * (pE + pW) + ay * (pN + pS) - breg[k]) * inv_diag;
      }
    }
    __syncthreads();
    T* t = src; src = dst; dst = t;
  }
#pragma unroll
  for (int k = 0; k < kCoarsePerThread; ++k) {
    const int c = threadIdx.x + k * kCoarseThreads;
    if (c < n) out[c] = src[c];
  }
}

// sum(r^2) partials (multigrid_solver.py:229-230), FP64 accumulation
template <typename T>
__global__ __launch_bounds__(256) void k_ref_resnorm(const T* __restrict__ p, const T* __restrict__ b,
                                                     int nx, int ny, T ax, T ay,
                                                     double* __restrict__ partial) {
  __shared__ double sh[16];
  const long n = (long)nx * ny;
  double acc = 0.0;
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long)gridDim.x * 256) {
    const int i = (int)(c / ny), j = (int)(c - (long)i * ny);
    const double r = (double)ref_residual(p, b, 1, nx, ny, ax, ay, i, j);
    acc += r * r;
  }
  const double t = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ void k_ref_resnorm_final(const double* __restrict__ partial, int np, double dxdy,
                                    double tol2, aj_mg_stats* st) {
  __shared__ double sh[16];
  double acc = 0.0;
  for (int k = threadIdx.x; k < np; k += 256) acc += partial[k];
  const double tot = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double nsq = tot * dxdy;
    const int executed = st->converged == 0;
    const int k = st->outer;  // evaluation index
    if (k < AJ_MG_MAX_NORMS) st->norms[k] = nsq;
    st->outer = k + 1;
    if (executed) st->cycles += 1;
    st->norm = nsq;
    if (!isfinite(nsq)) st->nonfinite = 1;
    st->converged = (nsq < tol2) ? 1 : 0;
  }
}

__global__ void k_mg_begin(aj_mg_stats* st, int* state) {
  if (threadIdx.x == 0) {
    st->cycles = 0;
    st->outer = 0;
    st->converged = 0;
    st->nonfinite = 0;
    st->norm = 0.0;
    if (state) state[0] = 0;
  }
  for (int k = threadIdx.x; k < AJ_MG_MAX_NORMS; k += blockDim.x) st->norms[k] = 0.0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_or_zero(const T* __restrict__ src, T* __restrict__ dst,
                                                      long n) {
  const long c = (long)blockIdx.x * 256 + threadIdx.x;
  if (c < n) dst[c] = src ? src[c] : T(0);
}

constexpr int kJX = 16, kJY = 64;

template <typename T>
static int ref_jacobi(const aj_mg_plan* pl, const T* p_in, const T* e_c, int ncx, int ncy,
                      const T* b, int bzl, T* p_out, int nx, int ny, int clamp, hipStream_t st) {
  RefJacobiArgs<T> a;
  const T dx = T(pl->dx), dy = T(pl->dy);
  a.p_in = p_in; a.e_c = e_c; a.ncx = ncx; a.ncy = ncy; a.b = b; a.b_zero_last = bzl;
  a.p_out = p_out; a.nx = nx; a.ny = ny;
  a.ax = T(1) / (dx * dx);
  a.ay = T(1) / (dy * dy);
  a.inv_diag = T(1) / (T(2) * (a.ax + a.ay));
  a.clamp_last = clamp;
  a.gate = pl->stats;
  dim3 grid(cdiv(ny, kJY), cdiv(nx, kJX));
  hipLaunchKernelGGL((k_ref_jacobi<T, kJX, kJY, 2>), grid, dim3(256), 0, st, a);
  AJ_LAUNCH_CHECK("k_ref_jacobi");
  return AJ_OK;
}

template <typename T>
static bool coarse_fits(int n) {
  return n <= kCoarseThreads * kCoarsePerThread && (size_t)n * 2 * sizeof(T) <= 65536;
}

// v_cycle (multigrid_solver.py:186-217): result of level `lev` lands in p_dst
template <typename T>
static int ref_vcycle(aj_mg_plan* pl, int lev, const T* p_src, T* p_dst, const T* b, int bzl,
                      hipStream_t st) {
  const int nx = pl->lnx[lev], ny = pl->lny[lev];
  const T dx = T(pl->dx), dy = T(pl->dy);
  const T ax = T(1) / (dx * dx), ay = T(1) / (dy * dy);
  const T inv_diag = T(1) / (T(2) * (ax + ay));
  if (lev >= pl->ml) {
    const int n = nx * ny;
    if (coarse_fits<T>(n)) {
      hipLaunchKernelGGL(k_ref_coarse<T>, dim3(1), dim3(kCoarseThreads), 2 * n * sizeof(T), st, b,
                         p_dst, nx, ny, ax, ay, inv_diag, 10, pl->stats);
      AJ_LAUNCH_CHECK("k_ref_coarse");
      return AJ_OK;
    }
    // 10 sweeps as 5 fused pairs, ping-pong ending in p_dst
    T* other = (T*)pl->Y[lev];
    const T* src = nullptr;
    for (int k = 0; k < 5; ++k) {
      T* dst = (k % 2 == 0) ? p_dst : other;
      if (int rc = ref_jacobi<T>(pl, src, nullptr, 0, 0, b, bzl, dst, nx, ny, 0, st)) return rc;
      src = dst;
    }
    return AJ_OK;
  }
  T* Y = (T*)pl->Y[lev];
  if (int rc = ref_jacobi<T>(pl, p_src, nullptr, 0, 0, b, bzl, Y, nx, ny, 1, st)) return rc;
  const int ncx = pl->lnx[lev + 1], ncy = pl->lny[lev + 1];
  const int degenerate = (nx / 2 < 2 || ny / 2 < 2) ? 1 : 0;
  T* Bc = (T*)pl->B[lev + 1];
  hipLaunchKernelGGL(k_ref_resrestrict<T>, dim3(cdiv(ncy, 256), ncx), dim3(256), 0, st,
                     (const T*)Y, b, bzl, nx, ny, ax, ay, Bc, ncx, ncy, degenerate, pl->stats);
  AJ_LAUNCH_CHECK("k_ref_resrestrict");
  T* Ec = (T*)pl->E[lev + 1];
  if (int rc = ref_vcycle<T>(pl, lev + 1, nullptr, Ec, Bc, 0, st)) return rc;
  return ref_jacobi<T>(pl, Y, Ec, ncx, ncy, b, bzl, p_dst, nx, ny, 1, st);
}

template <typename T>
static int ref_solve(aj_mg_plan* pl, const T* rhs, const T* p0, T* p_out, int v_cycles, double tol,
                     hipStream_t st) {
  const int nx = pl->nx, ny = pl->ny;
  const long n = (long)nx * ny;
  hipLaunchKernelGGL(k_mg_begin, dim3(1), dim3(64), 0, st, pl->stats, pl->state);
  AJ_LAUNCH_CHECK("k_mg_begin");
  if (v_cycles <= 0) {
    hipLaunchKernelGGL(k_copy_or_zero<T>, dim3(cdiv(n, 256)), dim3(256), 0, st, p0, p_out, n);
    AJ_LAUNCH_CHECK("k_copy_or_zero");
    return AJ_OK;
  }
  const T dx = T(pl->dx), dy = T(pl->dy);
  const T ax = T(1) / (dx * dx), ay = T(1) / (dy * dy);
  const int nb = (int)std::min<long>(kNPartial, cdiv(n, 256));
  // tolerance**2 in the working precision (tolerance is a traced f32 scalar)
  const T tolT = T(tol);
  const double tol2 = (double)(tolT * tolT);
  const double dxdy = (double)dx * (double)dy;
  for (int c = 0; c < v_cycles; ++c) {
    const T* src = (c == 0) ? p0 : p_out;
    if (int rc = ref_vcycle<T>(pl, 0, src, p_out, rhs, 1, st)) return rc;
    hipLaunchKernelGGL(k_ref_resnorm<T>, dim3(nb), dim3(256), 0, st, (const T*)p_out, rhs, nx, ny,
                       ax, ay, pl->partial);
    AJ_LAUNCH_CHECK("k_ref_resnorm");
    hipLaunchKernelGGL(k_ref_resnorm_final, dim3(1), dim3(256), 0, st, (const double*)pl->partial,
                       nb, dxdy, tol2, pl->stats);
    AJ_LAUNCH_CHECK("k_ref_resnorm_final");
  }
  return AJ_OK;
}

// reference level hierarchy (multigrid_solver.py:96-106, 108-156)
static int ref_shapes(aj_mg_plan* p) {
  int ml = 1, tx = p->nx, ty = p->ny;
  while (tx % 2 == 0 && ty % 2 == 0 && ml < p->levels) {
    tx /= 2;
    ty /= 2;
    ++ml;
  }
  char msg[256];
  std::snprintf(msg, sizeof(msg),
                "Grid dimensions (%d, %d) not suitable for multigrid with %d levels. Grid must be "
                "divisible by 2^%d",
                p->nx, p->ny, p->levels, p->levels);
  if (ml < 2) return fail(AJ_ERR_GRID, msg);
  p->ml = ml;
  p->lnx.assign(1, p->nx);
  p->lny.assign(1, p->ny);
  for (int l = 0; l < ml; ++l) {
    const int sx = p->lnx.back(), sy = p->lny.back();
    int cx = sx / 2, cy = sy / 2;
    if (cx < 2 || cy < 2) cx = cy = 1;
    if (2 * cx != sx || 2 * cy != sy) return fail(AJ_ERR_GRID, msg);
    p->lnx.push_back(cx);
    p->lny.push_back(cy);
  }
  return AJ_OK;
}

struct Carver {
  char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    void* r = base ? base + off : nullptr;
    off += (bytes + 255) & ~size_t(255);
    return r;
  }
};

static int ref_create(aj_mg_plan* p) {
  if (int rc = ref_shapes(p)) return rc;
  const size_t es = p->dtype == AJ_F64 ? 8 : 4;
  const int L = p->ml + 1;
  for (int pass = 0; pass < 2; ++pass) {
    Carver cv{pass == 0 ? nullptr : (char*)p->ws};
    p->B.assign(L, nullptr);
    p->Y.assign(L, nullptr);
    p->E.assign(L, nullptr);
    for (int l = 0; l < L; ++l) {
      const size_t n = (size_t)p->lnx[l] * p->lny[l];
      if (l > 0) {
        p->B[l] = cv.take(n * es);
        p->E[l] = cv.take(n * es);
      }
      p->Y[l] = cv.take(n * es);
    }
    p->partial = (double*)cv.take(kNPartial * sizeof(double));
    p->npartial = kNPartial;
    p->stats = (aj_mg_stats*)cv.take(sizeof(aj_mg_stats));
    p->state = (int*)cv.take(64);
    if (pass == 0) {
      p->ws_bytes = cv.off;
      AJ_HIP(hipMalloc(&p->ws, p->ws_bytes));
    }
  }
  return AJ_OK;
}

}  // namespace aj

using namespace aj;

extern "C" int aj_mg_create(aj_mg_plan** out, int dtype, int variant, int nx, int ny, double dx,
                            double dy, int levels) {
  if (!out) return fail(AJ_ERR_INVALID, "aj_mg_create: null plan pointer");
  *out = nullptr;
  if (dtype != AJ_F32 && dtype != AJ_F64) return fail(AJ_ERR_INVALID, "aj_mg_create: bad dtype");
  if (variant != AJ_MG_REFERENCE && variant != AJ_MG_TEXTBOOK)
    return fail(AJ_ERR_INVALID, "aj_mg_create: bad variant");
  if (nx < 2 || ny < 2 || (long)nx * ny > (1L << 31))
    return fail(AJ_ERR_INVALID, "aj_mg_create: bad grid size");
  if (!(dx > 0) || !(dy > 0)) return fail(AJ_ERR_INVALID, "aj_mg_create: dx, dy must be > 0");
  aj_mg_plan* p = new aj_mg_plan();
  p->dtype = dtype;
  p->variant = variant;
  p->nx = nx;
  p->ny = ny;
  p->dx = dx;
  p->dy = dy;
  p->levels = levels;
  if (hipGetDevice(&p->device) != hipSuccess) p->device = 0;
  const int rc = (variant == AJ_MG_REFERENCE) ? ref_create(p) : tb_create(p);
  if (rc != AJ_OK) {
    if (p->ws) (void)hipFree(p->ws);
    delete p;
    return rc;
  }
  *out = p;
  return AJ_OK;
}

extern "C" int aj_mg_destroy(aj_mg_plan* p) {
  if (!p) return AJ_OK;
  for (int k = 0; k < 2; ++k) {
    if (p->graph_exec[k]) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec[k]);
    if (p->graph[k]) (void)hipGraphDestroy((hipGraph_t)p->graph[k]);
    p->graph_exec[k] = p->graph[k] = nullptr;
  }
  if (p->host_stats) {
    (void)hipHostFree(p->host_stats);
    p->host_stats = nullptr;
  }
  if (p->ws) {
    hipError_t e = hipFree(p->ws);
    p->ws = nullptr;
    delete p;
    if (e != hipSuccess) return hip_fail(e, "aj_mg_destroy");
    return AJ_OK;
  }
  delete p;
  return AJ_OK;
}

extern "C" int aj_mg_set_option(aj_mg_plan* p, int option, int value) {
  if (!p) return fail(AJ_ERR_INVALID, "aj_mg_set_option: null plan");
  switch (option) {
    case AJ_MG_OPT_ADAPTIVE: p->adaptive = value ? 1 : 0; return AJ_OK;
    case AJ_MG_OPT_GRAPH: p->use_graph = value ? 1 : 0; return AJ_OK;
    case AJ_MG_OPT_GUESS_TEST: p->guess_test = value ? 1 : 0; return AJ_OK;
    default: return fail(AJ_ERR_INVALID, "aj_mg_set_option: unknown option");
  }
}

extern "C" size_t aj_mg_workspace_bytes(const aj_mg_plan* p) { return p ? p->ws_bytes : 0; }

extern "C" int aj_mg_num_levels(const aj_mg_plan* p) { return p ? (int)p->lnx.size() : 0; }

extern "C" int aj_poisson_mg(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out,
                             int max_cycles, double tol, void* stream) {
  if (!p || !rhs || !p_out) return fail(AJ_ERR_INVALID, "aj_poisson_mg: null pointer");
  if (rhs == p_out || (p0 && p0 == p_out))
    return fail(AJ_ERR_INVALID, "aj_poisson_mg: p_out must not alias rhs or p0");
  hipStream_t st = (hipStream_t)stream;
  if (p->variant == AJ_MG_TEXTBOOK) return tb_solve(p, rhs, p0, p_out, max_cycles, tol, st);
  if (p->dtype == AJ_F32)
    return ref_solve<float>(p, (const float*)rhs, (const float*)p0, (float*)p_out, max_cycles, tol, st);
  return ref_solve<double>(p, (const double*)rhs, (const double*)p0, (double*)p_out, max_cycles, tol,
                           st);
}

extern "C" int aj_mg_get_stats(aj_mg_plan* p, aj_mg_stats* out, void* stream) {
  if (!p || !out) return fail(AJ_ERR_INVALID, "aj_mg_get_stats: null pointer");
  hipStream_t st = (hipStream_t)stream;
  AJ_HIP(hipMemcpyAsync(out, p->stats, sizeof(aj_mg_stats), hipMemcpyDeviceToHost, st));
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}
