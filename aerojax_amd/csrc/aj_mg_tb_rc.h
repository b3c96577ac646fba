// Register-column red-black smoothers of the textbook multigrid: further
// candidates of the smoother autotuner beside the LDS tiles, evaluating
// every cell with the same per-cell functions (tb_update, tb_res,
// tb_prolong_val, tb_weight), so that every candidate gives the same bits.
// Included by aj_mg_tb.hip.
//
// A workgroup of W waves owns a tile of W*R rows x 128 columns (n and m of
// every smoothed level are even).  Wave w holds rows gx0 + w*R .. + R - 1
// in registers and lane l the two columns gy0 + 2l, gy0 + 2l + 1 (gx0, gy0
// and R even), so on every row one of a lane's two cells is red and the
// other black, at a slot fixed at compile time: cell (row r, slot s) has
// colour (r + s) & 1.  A half-sweep updates one slot per row.  Its
// x-neighbours are the lane's own registers of rows r -+ 1 (across waves:
// the neighbouring wave's last or first row, exchanged through LDS after
// every half-sweep), its y-neighbours the lane's other slot and, by a DPP
// wave shift, the neighbouring lane's.  Nothing else goes through LDS, so a
// cell update is 6 VALU instructions (the LDS tiles: about four times that
// with the loads, masks and index arithmetic).
//
// Temporal blocking: the tile's outer rows and columns are halo owned by
// the neighbouring tiles; their values go stale from the tile edge inwards,
// one row or column per half-sweep.  Pre-smoothing (4 half-sweeps, then the
// residual) keeps 6 rows and 6 columns of halo on each side, post-smoothing
// 4 rows and 6 columns; the kept region starts at an even row and column,
// so every 2 x 2 restriction block lies in one wave and one lane.
//
// Walls: a lane whose columns lie outside [0, m) holds 0 and keeps it (its
// update weight is 0 and omw 1), which is the masked value of the absent
// y-neighbour of a wall cell.  Rows are uniform across a wave, so the
// inlet and outlet rows are handled by uniform branches, in a second
// instantiation of the tile body used only by tiles that reach them.
#pragma once

namespace rc {

constexpr int kLanes = 64;
constexpr int kCols = 2 * kLanes;  // columns of a wave's strip

// value of lane l - 1 in lane l (0 in lane 0)
__device__ __forceinline__ float from_left(float x) { return aj::dpp_from_left(x, 0.0f); }
// value of lane l + 1 in lane l (0 in lane 63)
__device__ __forceinline__ float from_right(float x) { return aj::dpp_from_right(x, 0.0f); }

// Tile of W waves x R rows with HX halo rows and HY halo columns per side.
template <int W, int R, int HX, int HY>
struct Geo {
  static_assert(R % 2 == 0 && HX % 2 == 0 && HY % 2 == 0, "even rows per wave and halo");
  static constexpr int ROWS = W * R;         // rows of a tile
  static constexpr int U = ROWS - 2 * HX;    // kept rows
  static constexpr int UY = kCols - 2 * HY;  // kept columns
  static_assert(U >= 2, "tile shorter than its halo");
  // x-tiles with no halo on the inlet and outlet sides (edge_x_tiles)
  __host__ __device__ static int x_tiles(int n) { return aj::edge_x_tiles(n, ROWS, HX); }
  __host__ __device__ static void x_tile(int xt, int n, int* gx0, int* k0, int* k1) {
    aj::edge_x_tile(xt, n, ROWS, HX, gx0, k0, k1);
  }
};
// pre-smoothing: 4 half-sweeps and the residual spoil 5 rows / columns from
// an edge; post-smoothing: 4 half-sweeps spoil 4 rows, and 5 columns, since
// the prolongation of a strip's first and last column reads a coarse value
// across the strip edge
template <int W, int R> using PreGeo = Geo<W, R, 6, 6>;
template <int W, int R> using PostGeo = Geo<W, R, 4, 6>;

// Tile (xt, yt) of this workgroup among XT x YT (yt fastest), numbered so
// that the workgroups of one XCD take one contiguous range of tiles (the
// hardware deals workgroup ids to the 8 XCDs in turn; gridDim.x is a
// multiple of 8).  False for the padding workgroups past the last tile.
__device__ __forceinline__ bool tile_of(int XT, int YT, int* xt, int* yt) {
  const int T = XT * YT, per = (int)(gridDim.x >> 3);
  const int k = (int)blockIdx.x;
  const int t = (k & 7) * per + (k >> 3);
  if (t >= T) return false;
  *xt = t / YT;
  *yt = t - *xt * YT;
  return true;
}

// Per-row facts of a tile row (uniform across the wave).
struct RowInfo {
  bool act;   // updated (0 <= gx < rows carrying an equation)
  bool hasW;  // gx > 0
  bool last;  // gx == n - 1
};

// Four red-black half-sweeps over the wave's rows p[R][2].  xb: the LDS
// exchange rows, [buffer][wave][first row = 0, last row = 1][lane].  GEN:
// the tile reaches the inlet or outlet row (rows with absent x-neighbours or
// no update); otherwise every row is interior in x.
template <int W, int R, bool GEN>
__device__ __forceinline__ void sweeps(const aj::TbLevel& L, float (&p)[R][2],
                                       const float (&b)[R][2], const float (&wr)[R][2],
                                       float omw, const RowInfo (&ri)[R],
                                       float (*xb)[W][2][kLanes], int w, int lane) {
  const int wu = w > 0 ? w - 1 : w, wd = w < W - 1 ? w + 1 : w;
  // before the first half-sweep (colour 0): the colour-1 cells of the first
  // and last rows, which the neighbouring waves read as x-neighbours
  xb[0][w][0][lane] = p[0][1];
  xb[0][w][1][lane] = p[R - 1][0];
  __syncthreads();
#pragma unroll
  for (int k = 1; k <= 4; ++k) {
    const int c = (k - 1) & 1;  // colour updated now
    const int rb = c;           // buffer the previous exchange wrote
    const float above = xb[rb][wu][1][lane], below = xb[rb][wd][0][lane];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int s = (c + r) & 1;
      if (GEN && !ri[r].act) continue;
      const float Wv = r > 0 ? p[r - 1][s] : above;
      const float Ev = r < R - 1 ? p[r + 1][s] : below;
      const float xs = GEN ? ((ri[r].last ? 0.0f : Ev) + (ri[r].hasW ? Wv : 0.0f)) : Ev + Wv;
      float ys;
      if (s == 0) {
        const float S = from_left(p[r][1]);
        ys = p[r][1] + S;
      } else {
        const float N = from_right(p[r][0]);
        ys = N + p[r][0];
      }
      p[r][s] = aj::tb_update(L, xs, ys, b[r][s], p[r][s], wr[r][s], omw);
    }
    // the cells just updated in the first and last rows
    xb[k & 1][w][0][lane] = p[0][c];
    xb[k & 1][w][1][lane] = p[R - 1][(c + R - 1) & 1];
    __syncthreads();
  }
}

// Facts of a lane: its first column, whether its columns are inside the
// grid, whether slot 0 has a south and slot 1 a north neighbour.
struct LaneInfo {
  int gy;
  bool in;
  bool hasS0, hasN1;
};

__device__ __forceinline__ LaneInfo lane_info(int gy0, int lane, int m) {
  LaneInfo li;
  li.gy = gy0 + 2 * lane;
  li.in = li.gy >= 0 && li.gy <= m - 2;
  li.hasS0 = li.gy > 0;
  li.hasN1 = li.gy + 1 < m - 1;
  return li;
}

template <int R, bool GEN>
__device__ __forceinline__ void row_infos(const aj::TbLevel& L, int gxw, RowInfo (&ri)[R]) {
  const int active_rows = (L.level == 0) ? L.n - 1 : L.n;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gx = gxw + r;
    ri[r].act = GEN ? (gx >= 0 && gx < active_rows) : true;
    ri[r].hasW = GEN ? gx > 0 : true;
    ri[r].last = GEN ? gx == L.n - 1 : false;
  }
}

// update weights per row and slot (0 for lanes outside the grid)
template <int R, bool GEN>
__device__ __forceinline__ void weights(const aj::TbLevel& L, const LaneInfo& li,
                                        const RowInfo (&ri)[R], float (&wr)[R][2]) {
  if (!GEN) {
    const float w0 = li.in ? aj::tb_weight(L, true, false, li.hasS0, true) : 0.0f;
    const float w1 = li.in ? aj::tb_weight(L, true, false, true, li.hasN1) : 0.0f;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      wr[r][0] = w0;
      wr[r][1] = w1;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      wr[r][0] = li.in ? aj::tb_weight(L, ri[r].hasW, ri[r].last, li.hasS0, true) : 0.0f;
      wr[r][1] = li.in ? aj::tb_weight(L, ri[r].hasW, ri[r].last, true, li.hasN1) : 0.0f;
    }
  }
}

// ---------------------------------------------------------------------------
// pre: 2 RB sweeps + residual + restriction
// ---------------------------------------------------------------------------
template <int W, int R, bool GEN>
__device__ __forceinline__ void pre_body(const aj::TbLevel& L, float (&p)[R][2],
                                         const float (&b)[R][2], const LaneInfo& li, int gx0,
                                         int k0, int k1, int gy0, float* __restrict__ Y,
                                         float* __restrict__ Bc, float (*xb)[W][2][kLanes]) {
  using G = PreGeo<W, R>;
  const int w = aj::wave_id(), lane = (int)(threadIdx.x & 63);
  const int gxw = gx0 + w * R;
  const int n = L.n, m = L.m;
  RowInfo ri[R];
  row_infos<R, GEN>(L, gxw, ri);
  float wr[R][2];
  weights<R, GEN>(L, li, ri, wr);
  const float omw = li.in ? 1.0f - L.omega : 1.0f;
  sweeps<W, R, GEN>(L, p, b, wr, omw, ri, xb, w, lane);
  // residual of the smoothed iterate; row -1 and row R come from the last
  // two exchanges (slot 0 of row -1 is colour 1, updated last)
  const int wu = w > 0 ? w - 1 : w, wd = w < W - 1 ? w + 1 : w;
  const float up0 = xb[0][wu][1][lane], up1 = xb[1][wu][1][lane];
  const float dn0 = xb[1][wd][0][lane], dn1 = xb[0][wd][0][lane];
  float res[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const float l0 = from_left(p[r][1]), r1 = from_right(p[r][0]);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
      const float pc = p[r][s];
      float Wv = r > 0 ? p[r - 1][s] : (s == 0 ? up0 : up1);
      float Ev = r < R - 1 ? p[r + 1][s] : (s == 0 ? dn0 : dn1);
      float Sv, Nv;
      if (s == 0) {
        Sv = li.hasS0 ? l0 : pc;
        Nv = p[r][1];
      } else {
        Sv = p[r][0];
        Nv = li.hasN1 ? r1 : pc;
      }
      if (GEN) {
        if (!ri[r].hasW) Wv = pc;
        if (ri[r].last) Ev = L.gam * pc;
      }
      float rv = aj::tb_res(L, pc, b[r][s], Wv, Ev, Sv, Nv);
      if (GEN && L.level == 0 && ri[r].last) rv = 0.0f;
      res[r][s] = rv;
    }
  }
  // kept cells: Y and the 2 x 2 restriction of the residual
  const bool lkeep = li.in && 2 * lane >= 6 && 2 * lane < 6 + G::UY;
  const int mc = m >> 1;
#pragma unroll
  for (int r = 0; r < R; r += 2) {
    const int gx = gxw + r;  // even
    const bool rkeep = gx >= k0 && gx < k1;  // k0 even: pairs kept together
    if (rkeep && lkeep) {
      *reinterpret_cast<float2*>(Y + (long)gx * m + li.gy) = make_float2(p[r][0], p[r][1]);
      *reinterpret_cast<float2*>(Y + (long)(gx + 1) * m + li.gy) =
          make_float2(p[r + 1][0], p[r + 1][1]);
      Bc[(long)(gx >> 1) * mc + (li.gy >> 1)] =
          0.25f * ((res[r][0] + res[r + 1][0]) + (res[r][1] + res[r + 1][1]));
    }
  }
}

// ---------------------------------------------------------------------------
// post: Y + prolong(E_coarse), 2 RB sweeps
// ---------------------------------------------------------------------------
template <int W, int R, bool GEN>
__device__ __forceinline__ void post_body(const aj::TbLevel& L, float (&p)[R][2],
                                          const float (&b)[R][2], const LaneInfo& li, int gx0,
                                          int k0, int k1, float* __restrict__ E,
                                          float (*xb)[W][2][kLanes]) {
  using G = PostGeo<W, R>;
  const int w = aj::wave_id(), lane = (int)(threadIdx.x & 63);
  const int gxw = gx0 + w * R;
  const int n = L.n, m = L.m;
  RowInfo ri[R];
  row_infos<R, GEN>(L, gxw, ri);
  float wr[R][2];
  weights<R, GEN>(L, li, ri, wr);
  const float omw = li.in ? 1.0f - L.omega : 1.0f;
  sweeps<W, R, GEN>(L, p, b, wr, omw, ri, xb, w, lane);
  const bool lkeep = li.in && 2 * lane >= 6 && 2 * lane < 6 + G::UY;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gx = gxw + r;
    const bool rkeep = gx >= k0 && gx < k1;
    if (rkeep && lkeep)
      *reinterpret_cast<float2*>(E + (long)gx * m + li.gy) = make_float2(p[r][0], p[r][1]);
  }
}

}  // namespace rc

// Pre-smoothing kernel: p_in (null = zero start) -> Y, coarse rhs Bc; grid
// (8 * ceil(XT * YT / 8), 1, instances), W waves per workgroup.
template <int W, int R>
__global__ __launch_bounds__(W * 64) void k_tb_pre_rc(TbLevel L, int XT, int YT,
                                                      const float* __restrict__ p_in,
                                                      const float* __restrict__ b,
                                                      float* __restrict__ Y,
                                                      float* __restrict__ Bc,
                                                      const aj_mg_stats* gate) {
  using G = rc::PreGeo<W, R>;
  __shared__ float xb[2][W][2][rc::kLanes];
  int xt, yt;
  if (!rc::tile_of(XT, YT, &xt, &yt)) return;
  const int n = L.n, m = L.m;
  {
    const long z = blockIdx.z;
    gate += z;
    if (p_in) p_in += z * n * m;
    b += z * n * m;
    Y += z * n * m;
    Bc += z * (n >> 1) * (m >> 1);
  }
  const int gated = gate->converged;
  const int w = aj::wave_id(), lane = (int)(threadIdx.x & 63);
  int gx0, k0, k1;
  G::x_tile(xt, n, &gx0, &k0, &k1);
  const int gy0 = yt * G::UY - 6;
  const rc::LaneInfo li = rc::lane_info(gy0, lane, m);
  const int gyl = li.in ? li.gy : 0;  // loads of lanes outside the grid read column 0
  float p[R][2], bb[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long g = (long)clampi(gx0 + w * R + r, 0, n - 1) * m + gyl;
    const float2 vb = *reinterpret_cast<const float2*>(b + g);
    const float2 vp = p_in ? *reinterpret_cast<const float2*>(p_in + g) : make_float2(0.0f, 0.0f);
    bb[r][0] = li.in ? vb.x : 0.0f;
    bb[r][1] = li.in ? vb.y : 0.0f;
    p[r][0] = li.in ? vp.x : 0.0f;
    p[r][1] = li.in ? vp.y : 0.0f;
  }
  if (gated) return;
  if (gx0 >= 1 && gx0 + G::ROWS <= n - 1)
    rc::pre_body<W, R, false>(L, p, bb, li, gx0, k0, k1, gy0, Y, Bc, xb);
  else
    rc::pre_body<W, R, true>(L, p, bb, li, gx0, k0, k1, gy0, Y, Bc, xb);
}

// Post-smoothing kernel: Y + prolong(Ec) -> E; same grid.
template <int W, int R>
__global__ __launch_bounds__(W * 64) void k_tb_post_rc(TbLevel L, int XT, int YT,
                                                       const float* __restrict__ Yin,
                                                       const float* __restrict__ Ec, int nc,
                                                       int mc, float gamc,
                                                       const float* __restrict__ b,
                                                       float* __restrict__ E,
                                                       const aj_mg_stats* gate) {
  using G = rc::PostGeo<W, R>;
  constexpr int NQ = R / 2 + 2;  // coarse rows of a wave's rows and their x-neighbours
  __shared__ float xb[2][W][2][rc::kLanes];
  int xt, yt;
  if (!rc::tile_of(XT, YT, &xt, &yt)) return;
  const int n = L.n, m = L.m;
  {
    const long z = blockIdx.z;
    gate += z;
    Yin += z * n * m;
    Ec += z * nc * mc;
    b += z * n * m;
    E += z * n * m;
  }
  const int gated = gate->converged;
  const int w = aj::wave_id(), lane = (int)(threadIdx.x & 63);
  int gx0, k0, k1;
  G::x_tile(xt, n, &gx0, &k0, &k1);
  const int gy0 = yt * G::UY - 6;
  const int gxw = gx0 + w * R;  // even
  const rc::LaneInfo li = rc::lane_info(gy0, lane, m);
  const int gyl = li.in ? li.gy : 0;
  // the lane's coarse column and coarse rows I0 - 1 .. I0 + R/2 (I0 = gxw / 2)
  const int J = clampi(li.gy >> 1, 0, mc - 1);
  const int I0 = gxw >> 1;
  float ec[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) ec[q] = Ec[(long)clampi(I0 - 1 + q, 0, nc - 1) * mc + J];
  float yv[R][2], bb[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const long g = (long)clampi(gxw + r, 0, n - 1) * m + gyl;
    const float2 vb = *reinterpret_cast<const float2*>(b + g);
    const float2 vy = *reinterpret_cast<const float2*>(Yin + g);
    bb[r][0] = li.in ? vb.x : 0.0f;
    bb[r][1] = li.in ? vb.y : 0.0f;
    yv[r][0] = vy.x;
    yv[r][1] = vy.y;
  }
  if (gated) return;
  // start value: Y + prolongation of the coarse correction (not on the
  // level-0 outlet row); lanes outside the grid hold 0
  const bool hasJl = J > 0, hasJr = J < mc - 1;
  float p[R][2];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int gx = gxw + r;
    const int I = I0 + r / 2;
    const int q = r / 2 + 1;
    int qn = (r & 1) ? q + 1 : q - 1;
    float sx = 1.0f;
    const int In = (r & 1) ? I + 1 : I - 1;
    if (In < 0) qn = q;
    else if (In >= nc) { qn = q; sx = gamc; }
    const float own = ec[q], xo = ec[qn];
    const float lo = rc::from_left(own), ro = rc::from_right(own);
    const float lx = rc::from_left(xo), rx = rc::from_right(xo);
    const bool keep = L.level == 0 && gx == n - 1;
    {
      const float yn = hasJl ? lo : own, dn = (hasJl ? lx : xo) * sx;
      const float pv = keep ? yv[r][0] : yv[r][0] + aj::tb_prolong_val(own, xo * sx, yn, dn);
      p[r][0] = li.in ? pv : 0.0f;
    }
    {
      const float yn = hasJr ? ro : own, dn = (hasJr ? rx : xo) * sx;
      const float pv = keep ? yv[r][1] : yv[r][1] + aj::tb_prolong_val(own, xo * sx, yn, dn);
      p[r][1] = li.in ? pv : 0.0f;
    }
  }
  if (gx0 >= 1 && gx0 + G::ROWS <= n - 1)
    rc::post_body<W, R, false>(L, p, bb, li, gx0, k0, k1, E, xb);
  else
    rc::post_body<W, R, true>(L, p, bb, li, gx0, k0, k1, E, xb);
}
