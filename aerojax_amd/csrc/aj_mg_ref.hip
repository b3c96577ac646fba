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
#include "aj_mg_ref_kernels.h"

#include <cmath>
#include <cstring>

namespace aj {

constexpr int kNPartial = 1024;

template <typename T, int TX, int TY, int NS>
__global__ __launch_bounds__(256) void k_ref_jacobi(RefJacobiArgs<T> a) {
  __shared__ JacobiLds<T, TX, TY, NS> L;
  ref_jacobi_tile<T, TX, TY, NS>(a, blockIdx.y * TX, blockIdx.x * TY, L);
}

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
  ref_resrestrict_cell(p, b, bzl, nx, ny, ax, ay, bc, ncy, degenerate, I, J);
}

template <typename T>
__global__ __launch_bounds__(kCoarseThreads) void k_ref_coarse(const T* __restrict__ b,
                                                               T* __restrict__ out, int nx, int ny,
                                                               T ax, T ay, T inv_diag, int nsw,
                                                               const aj_mg_stats* gate) {
  if (gated(gate)) return;
  extern __shared__ unsigned char smem_raw[];
  ref_coarse_wg(b, out, nx, ny, ax, ay, inv_diag, nsw, reinterpret_cast<T*>(smem_raw));
}

// sum(r^2) partials (multigrid_solver.py:229-230), FP64 accumulation
template <typename T>
__global__ __launch_bounds__(256) void k_ref_resnorm(const T* __restrict__ p, const T* __restrict__ b,
                                                     int nx, int ny, T ax, T ay,
                                                     double* __restrict__ partial) {
  __shared__ double sh[16];
  // grid-stride over the cells; (row, column) stepped without a division
  const int c0 = (int)(blockIdx.x * 256 + threadIdx.x), stride = (int)(gridDim.x * 256);
  const int di = stride / ny, dj = stride - di * ny;
  int i = c0 / ny, j = c0 - (c0 / ny) * ny;
  double acc = 0.0;
  for (; i < nx;) {
    const double r = (double)ref_residual(p, b, 1, nx, ny, ax, ay, i, j);
    acc += r * r;
    i += di;
    j += dj;
    if (j >= ny) {
      j -= ny;
      ++i;
    }
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

static int mg_create(aj_mg_plan** out, int dtype, int variant, int nx, int ny, double dx,
                     double dy, int levels, int count) {
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
  p->batch = count;
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

extern "C" int aj_mg_create(aj_mg_plan** out, int dtype, int variant, int nx, int ny, double dx,
                            double dy, int levels) {
  return mg_create(out, dtype, variant, nx, ny, dx, dy, levels, 1);
}

extern "C" int aj_mg_create_batched(aj_mg_plan** out, int dtype, int nx, int ny, double dx,
                                    double dy, int levels, int count) {
  if (out) *out = nullptr;
  if (count < 1 || count > 65535)
    return fail(AJ_ERR_INVALID, "aj_mg_create_batched: count must be in 1..65535");
  return mg_create(out, dtype, AJ_MG_TEXTBOOK, nx, ny, dx, dy, levels, count);
}

extern "C" int aj_mg_batch_count(const aj_mg_plan* p) { return p ? p->batch : 0; }

extern "C" int aj_mg_destroy(aj_mg_plan* p) {
  if (!p) return AJ_OK;
  // an open deferred solve and graph replays may still run on the plan's
  // buffers: wait for the plan's device before tearing anything down
  {
    int cur = 0;
    const bool have = hipGetDevice(&cur) == hipSuccess;
    if (hipSetDevice(p->device) == hipSuccess) {
      if (p->pend_active && p->pend_ev) (void)hipEventSynchronize(p->pend_ev);
      (void)hipDeviceSynchronize();
    }
    p->pend_active = 0;
    if (have) (void)hipSetDevice(cur);
  }
  for (int k = 0; k < 2; ++k) {
    if (p->graph_exec[k]) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec[k]);
    if (p->graph[k]) (void)hipGraphDestroy((hipGraph_t)p->graph[k]);
    p->graph_exec[k] = p->graph[k] = nullptr;
  }
  if (p->pend_ev) {
    (void)hipEventDestroy(p->pend_ev);
    p->pend_ev = nullptr;
  }
  if (p->hist[0]) {
    (void)hipFree(p->hist[0]);
    p->hist[0] = p->hist[1] = nullptr;
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
    case AJ_MG_OPT_RELATIVE: p->relative = value ? 1 : 0; return AJ_OK;
    case AJ_MG_OPT_OMEGA: {
      if (p->variant != AJ_MG_TEXTBOOK)
        return fail(AJ_ERR_UNSUPPORTED, "aj_mg_set_option: omega applies to textbook plans");
      if (value < 500 || value > 1900)
        return fail(AJ_ERR_INVALID, "aj_mg_set_option: omega (in thousandths) must be in 500..1900");
      if (p->pend_active)
        return fail(AJ_ERR_INVALID, "aj_mg_set_option: a deferred solve is open");
      // the captured outer iterations hold the old factor: capture again
      AJ_HIP(hipDeviceSynchronize());
      for (int k = 0; k < 2; ++k) {
        if (p->graph_exec[k]) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec[k]);
        if (p->graph[k]) (void)hipGraphDestroy((hipGraph_t)p->graph[k]);
        p->graph_exec[k] = p->graph[k] = nullptr;
      }
      p->omega = value / 1000.0;
      return AJ_OK;
    }
    case AJ_MG_OPT_CYCLES: {
      if (p->variant != AJ_MG_TEXTBOOK)
        return fail(AJ_ERR_UNSUPPORTED, "aj_mg_set_option: cycles per outer iteration apply to textbook plans");
      if (value < 1 || value > 8)
        return fail(AJ_ERR_INVALID, "aj_mg_set_option: V-cycles per outer iteration must be in 1..8");
      if (p->pend_active)
        return fail(AJ_ERR_INVALID, "aj_mg_set_option: a deferred solve is open");
      // the captured outer iterations hold the old cycle count: capture again
      AJ_HIP(hipDeviceSynchronize());
      for (int k = 0; k < 2; ++k) {
        if (p->graph_exec[k]) (void)hipGraphExecDestroy((hipGraphExec_t)p->graph_exec[k]);
        if (p->graph[k]) (void)hipGraphDestroy((hipGraph_t)p->graph[k]);
        p->graph_exec[k] = p->graph[k] = nullptr;
      }
      // the predicted iteration count of the next solve keeps its V-cycles
      p->last_outer = std::max(1, (p->last_outer * p->cycles_per_outer + value - 1) / value);
      p->cycles_per_outer = value;
      return AJ_OK;
    }
    case AJ_MG_OPT_TEMPORAL: {
      if (p->variant != AJ_MG_TEXTBOOK)
        return fail(AJ_ERR_UNSUPPORTED, "aj_mg_set_option: the temporal guess needs a textbook plan");
      if (p->pend_active)
        return fail(AJ_ERR_INVALID, "aj_mg_set_option: a deferred solve is open");
      if (value && !p->hist[0]) {
        const size_t n = (size_t)p->batch * p->nx * p->ny;
        void* mem = nullptr;
        AJ_HIP(hipMalloc(&mem, 2 * n * sizeof(double)));
        p->hist[0] = (double*)mem;
        p->hist[1] = p->hist[0] + n;
      }
      // (re)enabling starts a new history: the next solve has no guess
      p->temporal = value ? 1 : 0;
      p->hist_valid = 0;
      p->hist_count = 0;
      p->hist_last = 0;
      return AJ_OK;
    }
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

extern "C" int aj_poisson_mg_enqueue(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out,
                                     int max_cycles, double tol, void* stream) {
  if (!p || !rhs || !p_out) return fail(AJ_ERR_INVALID, "aj_poisson_mg_enqueue: null pointer");
  if (rhs == p_out || (p0 && p0 == p_out))
    return fail(AJ_ERR_INVALID, "aj_poisson_mg_enqueue: p_out must not alias rhs or p0");
  if (p->variant != AJ_MG_TEXTBOOK)
    return aj_poisson_mg(p, rhs, p0, p_out, max_cycles, tol, stream);
  if (!p->adaptive)  // a non-adaptive solve is fully enqueued already
    return tb_solve(p, rhs, p0, p_out, max_cycles, tol, (hipStream_t)stream);
  return tb_solve_enqueue(p, rhs, p0, p_out, max_cycles, tol, (hipStream_t)stream);
}

extern "C" int aj_poisson_mg_complete(aj_mg_plan* p, int* more) {
  if (!p || !more) return fail(AJ_ERR_INVALID, "aj_poisson_mg_complete: null pointer");
  if (p->variant != AJ_MG_TEXTBOOK) {
    *more = 0;
    return AJ_OK;
  }
  return tb_solve_complete(p, more);
}

extern "C" int aj_mg_get_stats(aj_mg_plan* p, aj_mg_stats* out, void* stream) {
  if (!p || !out) return fail(AJ_ERR_INVALID, "aj_mg_get_stats: null pointer");
  if (p->pend_active)
    return fail(AJ_ERR_INVALID,
                "aj_mg_get_stats: a deferred solve is open (call aj_poisson_mg_complete first)");
  hipStream_t st = (hipStream_t)stream;
  AJ_HIP(hipMemcpyAsync(out, p->stats, sizeof(aj_mg_stats), hipMemcpyDeviceToHost, st));
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}

extern "C" int aj_mg_get_stats_batched(aj_mg_plan* p, aj_mg_stats* out, void* stream) {
  if (!p || !out) return fail(AJ_ERR_INVALID, "aj_mg_get_stats_batched: null pointer");
  if (p->pend_active)
    return fail(AJ_ERR_INVALID,
                "aj_mg_get_stats_batched: a deferred solve is open (call aj_poisson_mg_complete "
                "first)");
  hipStream_t st = (hipStream_t)stream;
  AJ_HIP(hipMemcpyAsync(out, p->stats, sizeof(aj_mg_stats) * p->batch, hipMemcpyDeviceToHost, st));
  AJ_HIP(hipStreamSynchronize(st));
  return AJ_OK;
}
