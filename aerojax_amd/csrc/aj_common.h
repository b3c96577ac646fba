// Shared host/device helpers for the aerojax HIP library (gfx950 / CDNA4).
//
// Layout convention of every field: (nx, ny) row-major, y contiguous
// (index = i*ny + j); row i = 0 is the inlet, i = nx-1 the outlet,
// columns j = 0 and j = ny-1 the walls.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <functional>
#include <string>

#include "../../include/aerojax.h"

namespace aj {

// ---- error state --------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define AJ_HIP(call)                                        \
  do {                                                      \
    hipError_t aj_e_ = (call);                              \
    if (aj_e_ != hipSuccess) return ::aj::hip_fail(aj_e_, #call); \
  } while (0)

#define AJ_LAUNCH_CHECK(what)                               \
  do {                                                      \
    hipError_t aj_e_ = hipGetLastError();                   \
    if (aj_e_ != hipSuccess) return ::aj::hip_fail(aj_e_, what); \
  } while (0)

inline unsigned cdiv(long a, long b) { return (unsigned)((a + b - 1) / b); }

// ---- live kernel timing (HIP events around selected launches) -----------
// Enabled by aj_profile_begin; every instrumented launch records an event
// pair on its stream and the algorithmic bytes it moves.
bool prof_on();
int prof_begin_launch(int tag, hipStream_t st);  // returns a slot or -1
void prof_end_launch(int slot, double bytes, hipStream_t st);
// launches a tuned kernel makes inside the open instrumented launch (the
// repeat factor while profiling, else 1), and how many it made
int prof_repeat();
void prof_count(int launches);

struct ProfScope {
  int slot;
  double bytes;
  hipStream_t st;
  ProfScope(int tag, double b, hipStream_t s) : slot(-1), bytes(b), st(s) {
    if (prof_on()) slot = prof_begin_launch(tag, s);
  }
  ~ProfScope() {
    if (slot >= 0) prof_end_launch(slot, bytes, st);
  }
};

// ---- launch-shape autotuner (aj_tune.hip) --------------------------------
// launch(c, stream) launches candidate c (returns an AJ status).  The first
// call for `key` times every candidate and keeps the fastest; later calls
// launch the kept candidate.  *pick receives the candidate used.
int tune_pick(const std::string& key, int ncand,
              const std::function<int(int, hipStream_t)>& launch, hipStream_t st, int* pick);

// ---- device helpers -----------------------------------------------------
// x-tiles of `rows` rows with `hx` halo rows per internal side over a grid
// of n rows: the first keeps rows [0, rows - hx) and the last up to
// rows - hx rows (the inlet and outlet rows have no neighbour beyond them,
// so those tiles need no halo on that side), the others rows - 2 hx; with
// rows and hx even every kept range starts on an even row.
__host__ __device__ inline int edge_x_tiles(int n, int rows, int hx) {
  const int u = rows - 2 * hx, f = rows - hx;
  return n <= rows ? 1 : 2 + (n - 2 * f > 0 ? (n - 2 * f + u - 1) / u : 0);
}
// origin gx0 and kept rows [k0, k1) of x-tile xt
__host__ __device__ inline void edge_x_tile(int xt, int n, int rows, int hx, int* gx0, int* k0,
                                            int* k1) {
  const int u = rows - 2 * hx, f = rows - hx;
  if (xt == 0) {
    *gx0 = 0;
    *k0 = 0;
    *k1 = n <= rows ? n : f;
  } else {
    *k0 = f + (xt - 1) * u;
    *gx0 = *k0 - hx;
    *k1 = *gx0 + rows >= n ? n : *k0 + u;
  }
}

__device__ __forceinline__ int clampi(int x, int lo, int hi) {
  return max(lo, min(x, hi));
}

// Index of the calling thread's wave in its workgroup, as a wave-uniform
// (scalar) value, so that row offsets and row tests derived from it stay
// in scalar registers.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Value of x in lane l - 1 of the wave, read by lane l (DPP wave_shr:1);
// lane 0, which has no such lane, gets fb.
__device__ __forceinline__ float dpp_from_left(float x, float fb) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(fb), __float_as_int(x), 0x138, 0xF, 0xF, false));
}
// Value of x in lane l + 1, read by lane l (DPP wave_shl:1); lane 63 gets fb.
__device__ __forceinline__ float dpp_from_right(float x, float fb) {
  return __int_as_float(
      __builtin_amdgcn_update_dpp(__float_as_int(fb), __float_as_int(x), 0x130, 0xF, 0xF, false));
}

// The hardware deals consecutive workgroup ids out to the 8 XCDs in turn;
// renumbering so that each XCD gets one contiguous range of blocks keeps
// the +-row neighbours of a block in the same L2.
__device__ __forceinline__ long xcd_block(long b, long nb) {
  if (nb % 8) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// (x, y) block coordinates of a 2D tile grid, numbered so that each XCD
// works on one contiguous band of tile rows (blockIdx.y): the halo rows a
// tile shares with the tiles above and below are then read into one L2.
__device__ __forceinline__ void tile_block(int* bx, int* by) {
  const long nb = (long)gridDim.x * gridDim.y;
  const long b = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, nb);
  *by = (int)(b / gridDim.x);
  *bx = (int)(b - (long)*by * gridDim.x);
}

// Row i and first column j0 of a block of a row kernel (grid
// (cdiv(ny, blockDim.x), nx), one row segment per block), numbered so that
// each XCD works on one contiguous band of rows.
__device__ __forceinline__ void row_block(int* i, int* j0) {
  const long nb = (long)gridDim.x * gridDim.y;
  const long b = xcd_block((long)blockIdx.y * gridDim.x + blockIdx.x, nb);
  *i = (int)(b / gridDim.x);
  *j0 = (int)(b - (long)*i * gridDim.x) * (int)blockDim.x;
}

// Sum of a double over the 64 lanes of a wave (fixed butterfly, every lane
// gets the same bits).
__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Block-wide sum of a double (blockDim.x multiple of 64, <= 1024) with a
// fixed reduction tree, so repeated runs give identical bits.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double t = 0.0;
  if (threadIdx.x < 64) {
    t = (lane < nw) ? sh[lane] : 0.0;
    for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  }
  __syncthreads();
  return t;  // valid in threads 0..63
}

// Block-wide maximum of non-negative doubles (same shape rules).
__device__ __forceinline__ double block_max(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  const int nw = blockDim.x >> 6;
  double t = 0.0;
  if (threadIdx.x < 64) {
    t = (lane < nw) ? sh[lane] : 0.0;
    for (int off = 32; off > 0; off >>= 1) t = fmax(t, __shfl_xor(t, off, 64));
  }
  __syncthreads();
  return t;  // valid in threads 0..63
}

}  // namespace aj
