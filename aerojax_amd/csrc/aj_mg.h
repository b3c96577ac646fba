// Multigrid plan shared by the reference-exact and the textbook solvers.
#pragma once

#include <vector>

#include "aj_common.h"

struct aj_mg_plan {
  int dtype = AJ_F32;
  int variant = AJ_MG_REFERENCE;
  int nx = 0, ny = 0;
  int levels = 4;     // requested levels (textbook: 0 = automatic)
  int batch = 1;      // textbook: independent instances of the grid, stored one after the other
  int ml = 0;         // reference: recursion depth (coarsest level index)
  double dx = 0, dy = 0;
  std::vector<int> lnx, lny;  // level shapes, level 0 = fine grid
  int device = 0;

  // one allocation holds every buffer below
  void* ws = nullptr;
  size_t ws_bytes = 0;

  // reference variant, per level (index = level); level 0 uses the
  // caller's rhs as b and the caller's output as E
  std::vector<void*> B, Y, E;

  // textbook variant
  std::vector<float*> tB, tY, tE;   // FP32 V-cycle buffers per level
  double* p64[2] = {nullptr, nullptr};  // FP64 iterate (device-selected ping-pong)
  double* coarse_C = nullptr;       // coarsest level: orthonormal DCT-II [m][m]
  float* coarse_Cf = nullptr;       // same, FP32
  double* coarse_coef = nullptr;    // coarsest level: PCR multipliers [steps][m][2][rows]
  double* coarse_dinv = nullptr;    // coarsest level: 1 / reduced diagonal [m][rows]
  double* coarse_W = nullptr;       // coarsest level: transformed rhs [m][rows]
  double* coarse_X = nullptr;       // coarsest level: per-mode solution [m][rows]
  int coarse_rows = 0;              // unknown rows of the coarsest level
  int coarse_steps = 0;             // PCR steps (ceil log2 rows)

  // reductions and device-side state
  double* partial = nullptr;
  double* partial2 = nullptr;       // textbook: zero-guess residual partials
  int npartial = 0;
  aj_mg_stats* stats = nullptr;     // device copy, also the cycle gate
  int* state = nullptr;             // textbook: [0] = current p64 buffer
  aj_mg_stats* host_stats = nullptr;  // pinned copy read by the adaptive driver
  int adaptive = 1;                 // textbook: read convergence between iterations
  int last_outer = 4;               // textbook: iterations the previous solve needed
  int use_graph = 1;                // textbook: replay a captured outer iteration
  int guess_test = 1;               // textbook: keep p0 only if ||r(p0)|| <= ||b||
  int relative = 0;                 // textbook: tolerance relative to ||b||
  double omega = 1.2;               // textbook: over-relaxation of the red-black smoother
  int cycles_per_outer = 2;         // textbook: V-cycles per FP64 defect-correction iteration
  // textbook: temporal initial guess from the previous solves' FP64 results
  int temporal = 0;                 // option on
  int hist_valid = 0;               // the last solve's iterate is still in p64[]
  int hist_count = 0;               // entries in the ring (0..2)
  int hist_last = 0;                // ring slot of the newest entry
  double* hist[2] = {nullptr, nullptr};  // one allocation, 2 x batch x nx x ny
  void* graph[2] = {nullptr, nullptr};       // hipGraph_t of an even / odd outer iteration
  void* graph_exec[2] = {nullptr, nullptr};  // hipGraphExec_t
  double* dtol = nullptr;           // textbook: tolerance of the current solve
  void* bcopy = nullptr;            // textbook: plan-owned copy of the rhs
  // textbook: a solve enqueued without its convergence read (completed by
  // tb_solve_complete): iterations enqueued, their bound, output, stream
  int pend_active = 0, pend_done = 0, pend_max = 0;
  void* pend_out = nullptr;
  hipStream_t pend_st = nullptr;
  hipEvent_t pend_ev = nullptr;     // recorded after the enqueued output
};

namespace aj {
int tb_create(aj_mg_plan* p);
int tb_solve(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out, int max_cycles,
             double tol, hipStream_t st);
// one to-tolerance solve of several textbook plans on one stream (p0 may be
// null, or hold null entries), with one host synchronisation per round of
// convergence checks for all of them
int tb_solve_many(aj_mg_plan* const* plans, int count, const void* const* rhs,
                  const void* const* p0, void* const* p_out, int max_cycles, double tol,
                  hipStream_t st);
// the same solve with its convergence read deferred: enqueue as many outer
// iterations as the previous solve needed and the output; tb_solve_complete
// then reads the device state (one synchronisation) and, if the solve has
// not converged, enqueues further iterations and the output again, setting
// *more
int tb_solve_enqueue(aj_mg_plan* p, const void* rhs, const void* p0, void* p_out, int max_cycles,
                     double tol, hipStream_t st);
int tb_solve_complete(aj_mg_plan* p, int* more);
}  // namespace aj
