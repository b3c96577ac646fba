// C++/OpenMP restatement of the to-tolerance projection step (CPU baseline).
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and by the CPU-baseline leg of
// bench.py (through oracle/cpu_step.py), never by the product package.  It
// is the same algorithm as the NumPy oracle (oracle/reference_step.py and
// oracle/textbook_mg.py), written for all host cores so that the GPU is
// timed next to a multi-threaded CPU run of identical work:
//   rk_step_unified, RK3, first-order upwind + nu lap, implicit Brinkman
//     factor with chi = sigmoid(-sdf / (0.1 dx)), hard mask, von Karman BCs
//     (advection_schemes/rk3_simple_new.py:175-358)
//   predictor BCs, rhs = divergence_nonperiodic(u*, v*) mask / dt
//     (solver/pure_functions/step.py:182-199, solver/operators.py:81-96)
//   the textbook multigrid for the outlet-Dirichlet operator: FP32 V(2,2)
//     cycles (red-black Gauss-Seidel with the exact row solve over-relaxed
//     by omega = 1.2, 2x2 average
//     restriction, cell-centred bilinear prolongation, exact FP64 coarsest
//     solve by a cosine transform along y and one tridiagonal solve per
//     mode) inside an FP64 defect correction, 2 V-cycles per outer
//     iteration, until sqrt(sum_{i<nx-1} r^2 dx dy) <= tol
//   projection with the interface projection of grad p, dp/dx = 0 on the
//     inlet row, hard mask, slip BCs, hard mask (step.py:257-397).
// Fields are (nx, ny) row-major, y contiguous, FP32.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kCoarseCells = 65536;
constexpr int kMaxCoarseRows = 1024;
constexpr float kOmega = 1.2f;  // over-relaxation of the red-black smoother

double outlet_gamma(int level) { return 1.0 - 1.0 / (0.5 - std::ldexp(1.0, -level - 1)); }

struct Level {
  int n = 0, m = 0, lev = 0;
  float axl = 0, ayl = 0, gam = 0, dE_last = 0;
  std::vector<float> b, p, r;
};

struct Solver {
  int nx = 0, ny = 0;
  double dx = 0, dy = 0, ax = 0, ay = 0;
  std::vector<Level> L;
  // exact coarsest solve
  int rows = 0, cm = 0;
  double coff = 0;
  std::vector<double> C;     // [k][j] orthonormal DCT-II
  std::vector<double> cdiag;  // [k][i]
  std::vector<double> W, X, dd;
  // defect correction
  std::vector<double> b64, p64, r64;
  std::vector<float> e32;
  // step fields
  std::vector<float> k1u, k1v, su, sv, tu, tv, us, vs, rhs, chi;
};

// ---------------------------------------------------------------- multigrid

void residual32(const Level& l, const float* p, const float* b, float* r) {
  const int n = l.n, m = l.m;
  const float axl = l.axl, ayl = l.ayl, gam = l.gam;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < m; ++j) {
      const long c = (long)i * m + j;
      if (l.lev == 0 && i == n - 1) {
        r[c] = 0.0f;
        continue;
      }
      const float pc = p[c];
      const float W = i > 0 ? p[c - m] : pc;
      const float E = i < n - 1 ? p[c + m] : (l.lev == 0 ? pc : gam * pc);
      const float S = j > 0 ? p[c - 1] : pc;
      const float N = j < m - 1 ? p[c + 1] : pc;
      const float Ap = axl * ((E - pc) + (W - pc)) + ayl * ((N - pc) + (S - pc));
      r[c] = b[c] - Ap;
    }
  }
}

void rbgs(const Level& l, float* p, const float* b, int sweeps) {
  const int n = l.n, m = l.m;
  const float axl = l.axl, ayl = l.ayl;
  const int last = l.lev == 0 ? n - 1 : n;  // level 0: Dirichlet row stays
  for (int s = 0; s < sweeps; ++s) {
    for (int color = 0; color < 2; ++color) {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < last; ++i) {
        for (int j = (i + color) & 1; j < m; j += 2) {
          const long c = (long)i * m + j;
          const bool hW = i > 0, hS = j > 0, hN = j < m - 1;
          const float E = i < n - 1 ? p[c + m] : 0.0f;
          const float dE = (l.lev > 0 && i == n - 1) ? l.dE_last : axl;
          const float W = hW ? p[c - m] : 0.0f;
          const float S = hS ? p[c - 1] : 0.0f;
          const float N = hN ? p[c + 1] : 0.0f;
          const float diag = dE + (hW ? axl : 0.0f) + (hS ? ayl : 0.0f) + (hN ? ayl : 0.0f);
          const float off = axl * (E + W) + ayl * (N + S);
          const float gs = (off - b[c]) / diag;
          p[c] = p[c] + kOmega * (gs - p[c]);
        }
      }
    }
  }
}

void restrict_avg(const Level& f, const float* r, Level& c) {
  const int m = f.m, nc = c.n, mc = c.m;
#pragma omp parallel for schedule(static)
  for (int I = 0; I < nc; ++I) {
    for (int J = 0; J < mc; ++J) {
      const int i0 = 2 * I, j0 = 2 * J;
      auto rv = [&](int i, int j) {
        return (f.lev == 0 && i == f.n - 1) ? 0.0f : r[(long)i * m + j];
      };
      c.b[(long)I * mc + J] =
          0.25f * ((rv(i0, j0) + rv(i0 + 1, j0)) + (rv(i0, j0 + 1) + rv(i0 + 1, j0 + 1)));
    }
  }
}

void prolong_add(const Level& c, const float* e, Level& f) {
  const int nc = c.n, mc = c.m, nf = f.n, mf = f.m;
  const float sg = (float)outlet_gamma(f.lev + 1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nf; ++i) {
    if (f.lev == 0 && i == nf - 1) continue;  // Dirichlet row
    const int I = i / 2, In = (i % 2 == 0) ? I - 1 : I + 1;
    const float xs = In >= nc ? sg : 1.0f;
    const int Ic = std::min(std::max(In, 0), nc - 1);
    for (int j = 0; j < mf; ++j) {
      const int J = j / 2, Jn = (j % 2 == 0) ? J - 1 : J + 1;
      const int Jc = std::min(std::max(Jn, 0), mc - 1);
      const float own = e[(long)I * mc + J];
      const float xn = e[(long)Ic * mc + J] * xs;
      const float yn = e[(long)I * mc + Jc];
      const float dn = e[(long)Ic * mc + Jc] * xs;
      f.p[(long)i * mf + j] += 0.5625f * own + 0.1875f * (xn + yn) + 0.0625f * dn;
    }
  }
}

void coarse_solve(Solver& S, const Level& l, const float* b, float* out) {
  const int m = S.cm, rows = S.rows;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < rows; ++i)
    for (int k = 0; k < m; ++k) {
      double acc = 0.0;
      const double* Ck = &S.C[(size_t)k * m];
      const float* bi = b + (long)i * m;
      for (int j = 0; j < m; ++j) acc += (double)bi[j] * Ck[j];
      S.W[(size_t)k * rows + i] = acc;
    }
  const double off = S.coff;
#pragma omp parallel for schedule(static)
  for (int k = 0; k < m; ++k) {
    double* x = &S.W[(size_t)k * rows];
    double* d = &S.dd[(size_t)k * rows];
    const double* d0 = &S.cdiag[(size_t)k * rows];
    for (int i = 0; i < rows; ++i) d[i] = d0[i];
    for (int i = 1; i < rows; ++i) {
      const double f = off / d[i - 1];
      d[i] -= f * off;
      x[i] -= f * x[i - 1];
    }
    x[rows - 1] /= d[rows - 1];
    for (int i = rows - 2; i >= 0; --i) x[i] = (x[i] - off * x[i + 1]) / d[i];
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < l.n; ++i)
    for (int j = 0; j < m; ++j) {
      double acc = 0.0;
      if (i < rows)
        for (int k = 0; k < m; ++k) acc += S.W[(size_t)k * rows + i] * S.C[(size_t)k * m + j];
      out[(long)i * m + j] = (float)acc;
    }
}

// one V-cycle on level `lv` with rhs L[lv].b into L[lv].p (zero start if `zero`)
void vcycle(Solver& S, int lv, bool zero) {
  Level& l = S.L[lv];
  const size_t cells = (size_t)l.n * l.m;
  if (lv == (int)S.L.size() - 1) {
    coarse_solve(S, l, l.b.data(), l.p.data());
    return;
  }
  if (zero) std::fill(l.p.begin(), l.p.begin() + cells, 0.0f);
  rbgs(l, l.p.data(), l.b.data(), 2);
  residual32(l, l.p.data(), l.b.data(), l.r.data());
  Level& c = S.L[lv + 1];
  restrict_avg(l, l.r.data(), c);
  vcycle(S, lv + 1, true);
  prolong_add(c, c.p.data(), l);
  rbgs(l, l.p.data(), l.b.data(), 2);
}

double residual64(Solver& S) {
  const int n = S.nx, m = S.ny;
  const double ax = S.ax, ay = S.ay;
  double sum = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < m; ++j) {
      const long c = (long)i * m + j;
      if (i == n - 1) {
        S.r64[c] = 0.0;
        continue;
      }
      const double* p = S.p64.data();
      const double pc = p[c];
      const double W = i > 0 ? p[c - m] : pc;
      const double E = p[c + m];  // row n-1 is the Dirichlet value 0
      const double Sv = j > 0 ? p[c - 1] : pc;
      const double N = j < m - 1 ? p[c + 1] : pc;
      const double r = S.b64[c] - (ax * ((E - pc) + (W - pc)) + ay * ((N - pc) + (Sv - pc)));
      S.r64[c] = r;
      sum += r * r;
    }
  }
  return std::sqrt(sum * S.dx * S.dy);
}

// FP64 defect correction; returns V-cycles (coarse solves on a one-level
// hierarchy), or -1 if the tolerance was not reached
int solve(Solver& S, const float* rhs, float* p_out, double tol, int max_outer, double* norm) {
  const int n = S.nx, m = S.ny;
  const size_t cells = (size_t)n * m;
  for (size_t c = 0; c < cells; ++c) {
    S.b64[c] = (c / m == (size_t)n - 1) ? 0.0 : (double)rhs[c];
    S.p64[c] = 0.0;
  }
  Level& l0 = S.L[0];
  const int cpo = S.L.size() == 1 ? 1 : 2;
  int cycles = 0, outer = 0;
  double nr = residual64(S);
  while (nr > tol && outer < max_outer) {
    for (size_t c = 0; c < cells; ++c) l0.b[c] = (float)S.r64[c];
    for (int k = 0; k < cpo; ++k) {
      vcycle(S, 0, k == 0);
      ++cycles;
    }
#pragma omp parallel for schedule(static)
    for (long c = 0; c < (long)cells; ++c) S.p64[c] += (double)l0.p[c];
    ++outer;
    nr = residual64(S);
  }
  for (size_t c = 0; c < cells; ++c) p_out[c] = (float)S.p64[c];
  *norm = nr;
  return nr <= tol ? cycles : -1;
}

// ---------------------------------------------------------------- RK step

inline float upwind(const float* f, long c, int i, int j, int n, int m, float a, float b,
                    float dx, float dy) {
  const float fc = f[c];
  const float fW = i > 0 ? f[c - m] : fc, fE = i < n - 1 ? f[c + m] : fc;
  const float fS = j > 0 ? f[c - 1] : fc, fN = j < m - 1 ? f[c + 1] : fc;
  const float fx = a > 0 ? a * (fc - fW) / dx : a * (fE - fc) / dx;
  const float fy = b > 0 ? b * (fc - fS) / dy : b * (fN - fc) / dy;
  return fx + fy;
}

inline float lap(const float* f, long c, int i, int j, int n, int m, float dx2, float dy2) {
  const float fc = f[c];
  const float fW = i > 0 ? f[c - m] : fc, fE = i < n - 1 ? f[c + m] : fc;
  const float fS = j > 0 ? f[c - 1] : fc, fN = j < m - 1 ? f[c + 1] : fc;
  return (fE + fW - 2.0f * fc) / dx2 + (fN + fS - 2.0f * fc) / dy2;
}

void rk_rhs(const float* u, const float* v, float* ku, float* kv, int n, int m, float nu,
            float dx, float dy) {
  const float dx2 = dx * dx, dy2 = dy * dy;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      const long c = (long)i * m + j;
      const float a = u[c], b = v[c];
      ku[c] = -upwind(u, c, i, j, n, m, a, b, dx, dy) + nu * lap(u, c, i, j, n, m, dx2, dy2);
      kv[c] = -upwind(v, c, i, j, n, m, a, b, dx, dy) + nu * lap(v, c, i, j, n, m, dx2, dy2);
    }
}

void velocity_bcs(float* u, float* v, int n, int m, float U_inf) {
  for (int j = 0; j < m; ++j) {
    u[j] = U_inf;
    v[j] = 0.0f;
    u[(long)(n - 1) * m + j] = u[(long)(n - 2) * m + j];
    v[(long)(n - 1) * m + j] = v[(long)(n - 2) * m + j];
  }
  for (int i = 0; i < n; ++i) {
    float* ur = u + (long)i * m;
    float* vr = v + (long)i * m;
    ur[0] = ur[1];
    ur[m - 1] = ur[m - 2];
    vr[0] = 0.0f;
    vr[m - 1] = 0.0f;
  }
}

void rk3_step(Solver& S, const float* u, const float* v, const float* mask, float nu, float dt,
              float dx, float dy, float U_inf) {
  const int n = S.nx, m = S.ny;
  const long cells = (long)n * m;
  float* D = S.chi.data();  // holds 1 + dt chi / eta
  rk_rhs(u, v, S.k1u.data(), S.k1v.data(), n, m, nu, dx, dy);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < cells; ++c) {
    S.su[c] = (u[c] + dt * S.k1u[c]) / D[c];
    S.sv[c] = (v[c] + dt * S.k1v[c]) / D[c];
  }
  rk_rhs(S.su.data(), S.sv.data(), S.k1u.data(), S.k1v.data(), n, m, nu, dx, dy);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < cells; ++c) {
    S.tu[c] = (0.75f * u[c] + 0.25f * (u[c] + dt * S.k1u[c])) / D[c];
    S.tv[c] = (0.75f * v[c] + 0.25f * (v[c] + dt * S.k1v[c])) / D[c];
  }
  rk_rhs(S.tu.data(), S.tv.data(), S.k1u.data(), S.k1v.data(), n, m, nu, dx, dy);
  const float third = 1.0f / 3.0f, two3 = 2.0f / 3.0f;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < cells; ++c) {
    const float a = (third * u[c] + two3 * (u[c] + dt * S.k1u[c])) / D[c];
    const float b = (third * v[c] + two3 * (v[c] + dt * S.k1v[c])) / D[c];
    const bool fluid = mask[c] > 0.5f;
    S.us[c] = fluid ? a : 0.0f;
    S.vs[c] = fluid ? b : 0.0f;
  }
  velocity_bcs(S.us.data(), S.vs.data(), n, m, U_inf);
  // predictor BCs of step_pure (inlet + slip walls): already in place
}

void pressure_rhs(Solver& S, const float* mask, float dt, float dx, float dy) {
  const int n = S.nx, m = S.ny;
  const float* u = S.us.data();
  const float* v = S.vs.data();
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      const long c = (long)i * m + j;
      float d = 0.0f;
      if (i > 0 && i < n - 1 && j > 0 && j < m - 1)
        d = (u[c + m] - u[c - m]) / (2.0f * dx) + (v[c + 1] - v[c - 1]) / (2.0f * dy);
      S.rhs[c] = d * mask[c] / dt;
    }
}

inline float gx_np(const float* f, long c, int i, int n, int m, float dx) {
  if (i == 0) return (f[c + m] - f[c]) / dx;
  if (i == n - 1) return (f[c] - f[c - m]) / dx;
  return (f[c + m] - f[c - m]) / (2.0f * dx);
}

inline float gy_np(const float* f, long c, int j, int m, float dy) {
  if (j == 0) return (f[c + 1] - f[c]) / dy;
  if (j == m - 1) return (f[c] - f[c - 1]) / dy;
  return (f[c + 1] - f[c - 1]) / (2.0f * dy);
}

void project(Solver& S, const float* p, const float* mask, float* u_out, float* v_out, float dt,
             float dx, float dy, float U_inf) {
  const int n = S.nx, m = S.ny;
  const float thr = 0.1f / dx;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < m; ++j) {
      const long c = (long)i * m + j;
      float gx = gx_np(p, c, i, n, m, dx), gy = gy_np(p, c, j, m, dy);
      const float mx = gx_np(mask, c, i, n, m, dx), my = gy_np(mask, c, j, m, dy);
      const float gm = std::sqrt(mx * mx + my * my);
      if (gm > thr) {
        const float nxv = mx / (gm + 1e-8f), nyv = my / (gm + 1e-8f);
        const float dpn = gx * nxv + gy * nyv;
        gx = gx - dpn * nxv;
        gy = gy - dpn * nyv;
      }
      if (i == 0) gx = 0.0f;
      const bool fluid = mask[c] > 0.5f;
      u_out[c] = fluid ? S.us[c] - dt * gx : 0.0f;
      v_out[c] = fluid ? S.vs[c] - dt * gy : 0.0f;
    }
  velocity_bcs(u_out, v_out, n, m, U_inf);
  const long cells = (long)n * m;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < cells; ++c)
    if (!(mask[c] > 0.5f)) u_out[c] = v_out[c] = 0.0f;
}

}  // namespace

extern "C" {

int ajcpu_threads(void) { return omp_get_max_threads(); }

void* ajcpu_create(int nx, int ny, double dx, double dy) {
  if (nx < 4 || ny < 2) return nullptr;
  Solver* S = new Solver();
  S->nx = nx;
  S->ny = ny;
  S->dx = dx;
  S->dy = dy;
  S->ax = 1.0 / (dx * dx);
  S->ay = 1.0 / (dy * dy);
  int n = nx, m = ny;
  std::vector<std::pair<int, int>> shapes{{n, m}};
  while ((n * m > kCoarseCells || n > kMaxCoarseRows) && n % 2 == 0 && m % 2 == 0 && n >= 4 &&
         m >= 2) {
    n /= 2;
    m /= 2;
    shapes.push_back({n, m});
  }
  const int Lc = (int)shapes.size() - 1;
  for (int l = 0; l <= Lc; ++l) {
    Level lv;
    lv.n = shapes[l].first;
    lv.m = shapes[l].second;
    lv.lev = l;
    const double s = std::ldexp(1.0, 2 * l);
    lv.axl = (float)(S->ax / s);
    lv.ayl = (float)(S->ay / s);
    lv.gam = l > 0 ? (float)outlet_gamma(l) : 0.0f;
    lv.dE_last = l > 0 ? (float)(1.0 - outlet_gamma(l)) * lv.axl : lv.axl;
    const size_t cells = (size_t)lv.n * lv.m;
    lv.b.assign(cells, 0.0f);
    lv.p.assign(cells, 0.0f);
    lv.r.assign(cells, 0.0f);
    S->L.push_back(std::move(lv));
  }
  // exact coarsest solve tables (FP64)
  const Level& c = S->L.back();
  S->cm = c.m;
  S->rows = Lc == 0 ? c.n - 1 : c.n;
  const double s = std::ldexp(1.0, 2 * Lc);
  const double axl = S->ax / s, ayl = S->ay / s;
  S->coff = axl;
  S->C.resize((size_t)c.m * c.m);
  for (int k = 0; k < c.m; ++k)
    for (int j = 0; j < c.m; ++j)
      S->C[(size_t)k * c.m + j] = std::sqrt((k == 0 ? 1.0 : 2.0) / c.m) *
                                  std::cos(M_PI * k * (j + 0.5) / c.m);
  S->cdiag.resize((size_t)c.m * S->rows);
  for (int k = 0; k < c.m; ++k) {
    const double sk = std::sin(M_PI * k / (2.0 * c.m));
    const double mu = -4.0 * ayl * sk * sk;
    for (int i = 0; i < S->rows; ++i) {
      const double dW = i > 0 ? axl : 0.0;
      const double dE = (Lc > 0 && i == S->rows - 1) ? axl * (1.0 - outlet_gamma(Lc)) : axl;
      S->cdiag[(size_t)k * S->rows + i] = -(dW + dE) + mu;
    }
  }
  S->W.assign((size_t)c.m * S->rows, 0.0);
  S->dd.assign((size_t)c.m * S->rows, 0.0);
  const size_t cells = (size_t)nx * ny;
  S->b64.assign(cells, 0.0);
  S->p64.assign(cells, 0.0);
  S->r64.assign(cells, 0.0);
  for (auto* f : {&S->k1u, &S->k1v, &S->su, &S->sv, &S->tu, &S->tv, &S->us, &S->vs, &S->rhs,
                  &S->chi})
    f->assign(cells, 0.0f);
  return S;
}

void ajcpu_destroy(void* h) { delete (Solver*)h; }

int ajcpu_levels(void* h) { return h ? (int)((Solver*)h)->L.size() : 0; }

// Multigrid alone: p_out solves the outlet-Dirichlet problem for rhs to
// sqrt(sum r^2 dx dy) <= tol; returns V-cycles, -1 if not reached.
int ajcpu_solve(void* h, const float* rhs, float* p_out, double tol, int max_outer,
                double* norm) {
  return solve(*(Solver*)h, rhs, p_out, tol, max_outer, norm);
}

// One to-tolerance step of the 2048x512 workload's composition: u, v are
// read and replaced by the new velocity; u_star / v_star / rhs / p_out
// (any may be null except p_out) receive the intermediate fields.  Returns
// the V-cycles of the pressure solve, -1 if the tolerance was not reached.
int ajcpu_step(void* h, float* u, float* v, const float* mask, const float* sdf, double nu,
               double dt, double U_inf, double eta, double tol, int max_outer, float* p_out,
               float* u_star, float* v_star, float* rhs_out, double* norm) {
  Solver& S = *(Solver*)h;
  const long cells = (long)S.nx * S.ny;
  const float fdt = (float)dt, fdx = (float)S.dx, fdy = (float)S.dy;
  const float feta = (float)eta, scale = 0.1f * fdx;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < cells; ++c) {
    const float chi = 1.0f / (1.0f + std::exp(sdf[c] / scale));  // sigmoid(-sdf / (0.1 dx))
    S.chi[c] = 1.0f + fdt * chi / feta;
  }
  rk3_step(S, u, v, mask, (float)nu, fdt, fdx, fdy, (float)U_inf);
  pressure_rhs(S, mask, fdt, fdx, fdy);
  const int cyc = solve(S, S.rhs.data(), p_out, tol, max_outer, norm);
  project(S, p_out, mask, u, v, fdt, fdx, fdy, (float)U_inf);
  if (u_star) std::memcpy(u_star, S.us.data(), cells * sizeof(float));
  if (v_star) std::memcpy(v_star, S.vs.data(), cells * sizeof(float));
  if (rhs_out) std::memcpy(rhs_out, S.rhs.data(), cells * sizeof(float));
  return cyc;
}

}  // extern "C"
