// 2D TM (Ez, Hx, Hy) Yee step for gfx950: the reference's 2D solver
// (src/neurosim/em/fdtd.py:298-353) in one pass per step.
//
// One launch per step reads buffer A and writes buffer B (ping-pong), so no
// workgroup reads a value another one already replaced.  The reference's
// step is H update -> E update -> damping or edge copies -> sources ->
// conductors; every output cell (i, j) is a pure function of A:
//
//   * the edge copies (fdtd.py:256-296: rows 0 and nx-1 first, then columns
//     0 and ny-1) give cell (i, j) the freshly updated values of cell
//     (ri(i), rj(j)), with r the identity inside and, at the edges,
//     reflecting 0 -> 1, n-1 -> n-2, periodic 0 -> n-2, n-1 -> 1;
//   * E at (a, b) needs the new H at (a, b), (a-1, b) and (a, b-1): the
//     thread recomputes the two neighbour values from A (or carries / takes
//     them from the adjacent lane, see tm2d_step_kernel);
//   * damping, the row sources and the conductor mask act at (i, j).
//
// Each float32 operation is the reference's, in its order, and the build
// uses -ffp-contract=off: the result is bit-identical to the float32
// restatement.  A workgroup is 16 wavefronts stacked along y, each walking
// 4 consecutive rows of one 64-column strip: a 64 x 64 tile.
// HBM traffic per cell: Ez, Hx, Hy read once and written once (24 B), plus
// one damping float in the absorbing band and one conductor byte when
// walls exist.

#include <hip/hip_runtime.h>

#include <cstdlib>
#include <string>

#include "tm2d.h"

namespace tm2d {

thread_local std::string g_err;

static int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define TM2D_CHECK(expr)                                                   \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) return fail(TM2D_EHIP, hipGetErrorString(e_));   \
  } while (0)

constexpr int kTx = 64;   // lanes along x
constexpr int kTy = 16;   // wavefronts per workgroup, stacked along y
constexpr int kRows = 4;  // rows marched by each wavefront

struct StepArgs {
  int nx, ny, npml;
  float ch, ce;
  const float* damp;
  const uint8_t* cond;
  int n_src;
  const int* rows;
  const float* vals;  // the source values of this step, n_src floats
  // tile order: workgroup L -> tile t; gx tiles per tile row, run > 0 gives
  // each XCD runs of `run` x-consecutive tiles (see tile_of)
  int gx, ntiles, run;
};

constexpr int kXcds = 8;  // MI355X: 8 XCDs, workgroups dealt round-robin

// Tile of workgroup L.  With run > 0 every group of 8 * run consecutive
// workgroups covers 8 * run consecutive tiles and XCD (L % 8) takes the
// run-long stretch (L % 8) * run .. of them: with run = gx / 8 each XCD
// holds the same eighth of every tile row, so a tile's x neighbours and
// (one group later) its y neighbours are read from the XCD's own L2.
__device__ __forceinline__ int tile_of(const StepArgs& g, int L) {
  if (g.run <= 0) return L;
  const int grp = kXcds * g.run;
  return (L / grp) * grp + (L % kXcds) * g.run + (L / kXcds) % g.run;
}

template <int MODE>
__device__ __forceinline__ int remap(int i, int n) {
  if (MODE == TM2D_REFLECTING) return i == 0 ? 1 : (i == n - 1 ? n - 2 : i);
  if (MODE == TM2D_PERIODIC) return i == 0 ? n - 2 : (i == n - 1 ? 1 : i);
  return i;
}

// The new (Ez, Hx, Hy) of output cell (i, j), from buffer A, for any cell
// (edge copies, wrap and sliced ranges included).
template <int MODE>
__device__ __forceinline__ void cell_general(const StepArgs& g, const float* __restrict__ ez,
                                             const float* __restrict__ hx,
                                             const float* __restrict__ hy, int i, int j,
                                             float& ezn_out, float& hxn_out, float& hyn_out) {
  constexpr bool kPer = MODE == TM2D_PERIODIC;
  const int nx = g.nx, ny = g.ny;
  const int a = remap<MODE>(i, nx);
  // x / y neighbours of (a, b): wrapped (periodic) or clamped; a clamped one
  // is loaded but not used (the update it would feed is the identity there)
  const int ap = kPer ? (a + 1 == nx ? 0 : a + 1) : min(a + 1, nx - 1);
  const int am = kPer ? (a == 0 ? nx - 1 : a - 1) : max(a - 1, 0);
  const int b = remap<MODE>(j, ny);
  const int bp = kPer ? (b + 1 == ny ? 0 : b + 1) : min(b + 1, ny - 1);
  const int bm = kPer ? (b == 0 ? ny - 1 : b - 1) : max(b - 1, 0);
  const int c = b * nx + a;
  const float e0 = ez[c];
  const float e_xp = ez[b * nx + ap];
  const float e_yp = ez[bp * nx + a];
  const float hy0 = hy[c];
  const float hx0 = hx[c];
  // fdtd.py:308-309 / 316-325: the new H at (a, b)
  const float hyn = (kPer || a <= nx - 2) ? hy0 + g.ch * (e_xp - e0) : hy0;
  const float hxn = (kPer || b <= ny - 2) ? hx0 - g.ch * (e_yp - e0) : hx0;
  float ezn = e0;
  if (kPer || (a >= 1 && b >= 1)) {
    // the new H at (a-1, b) and (a, b-1), recomputed from A
    const float e_xm = ez[b * nx + am];
    const float e_ym = ez[bm * nx + a];
    const float hyn_m = hy[b * nx + am] + g.ch * (e0 - e_xm);
    const float hxn_m = hx[bm * nx + a] - g.ch * (e0 - e_ym);
    // fdtd.py:311-313 / 328-335
    ezn = e0 + g.ce * ((hyn - hyn_m) - (hxn - hxn_m));
  }
  ezn_out = ezn;
  hxn_out = hxn;
  hyn_out = hyn;
}

// Damping, row sources and conductors at output cell (i, j), then the store
// (`conductor`: the cell's conductor flag, already loaded).
template <int MODE>
__device__ __forceinline__ void finish_cell_c(const StepArgs& g, int i, int j, bool in_x_band,
                                              bool conductor, float ezn, float hxn, float hyn,
                                              float* __restrict__ ez_out,
                                              float* __restrict__ hx_out,
                                              float* __restrict__ hy_out) {
  const int o = j * g.nx + i;
  if (MODE == TM2D_ABSORBING) {
    // fdtd.py:337-339; outside the band the factor is exp(-0) = 1
    if (in_x_band || j < g.npml || j >= g.ny - g.npml) ezn = ezn * g.damp[o];
  }
  // fdtd.py:249-254, 349: row sources in insertion order
  for (int s = 0; s < g.n_src; ++s)
    if (g.rows[s] == j) ezn = ezn + g.vals[s];
  // fdtd.py:351-352
  if (conductor) ezn = 0.0f;
  // non-temporal: outputs are read again only by the next step, after the
  // whole grid has streamed through the caches (16384^2 absorbing grid:
  // 1.272 -> 1.237 ms per step, the same bits)
  __builtin_nontemporal_store(ezn, &ez_out[o]);
  __builtin_nontemporal_store(hxn, &hx_out[o]);
  __builtin_nontemporal_store(hyn, &hy_out[o]);
}

template <int MODE>
__device__ __forceinline__ void finish_cell(const StepArgs& g, int i, int j, bool in_x_band,
                                            float ezn, float hxn, float hyn,
                                            float* __restrict__ ez_out,
                                            float* __restrict__ hx_out,
                                            float* __restrict__ hy_out) {
  const bool conductor = g.cond != nullptr && g.cond[j * g.nx + i];
  finish_cell_c<MODE>(g, i, j, in_x_band, conductor, ezn, hxn, hyn, ez_out, hx_out, hy_out);
}

// One workgroup: kTy wavefronts stacked along y, each a 64-column strip of
// kRows rows, all of whose loads are issued before the march (a 64 x 64
// tile; measured on the 16384^2 grid: 16 x 4 rows 1.288 ms per step,
// 8 x 8 1.324, 8 x 4 1.316, 4 x 8 1.424, 4 x 16 1.507, and 4 x 16 with
// the loads inside the march 1.461; the rows a wavefront shares with its
// neighbours are then read from cache).  A
// wavefront whose 64 columns all exist and whose rows all lie in
// [1, ny-2] (no edge copies, no wrap in y) marches down its rows carrying
// Ez(y-1), Ez(y) and the new Hx(y-1) in registers and taking the x
// neighbours Ez(x+-1), Hy(x-1) from the adjacent lanes (their loads are the
// same addresses); it loads Ez(y+1), Hx and Hy once per row.  Lanes whose
// x neighbour is not the adjacent lane's column (the wavefront's first and
// last lane, the edge-copy columns) load it.  Every other wavefront runs
// the general per-cell path.  Both paths perform the same float32
// operations in the same order.
template <int MODE>
__global__ __launch_bounds__(kTx * kTy) void tm2d_step_kernel(
    StepArgs g, const float* __restrict__ ez, const float* __restrict__ hx,
    const float* __restrict__ hy, float* __restrict__ ez_out, float* __restrict__ hx_out,
    float* __restrict__ hy_out) {
  constexpr bool kPer = MODE == TM2D_PERIODIC;
  const int nx = g.nx, ny = g.ny;
  const int lane = threadIdx.x;
  const int t = tile_of(g, (int)blockIdx.x);
  if (t >= g.ntiles) return;  // padding workgroups of the tile order
  const int x0 = (t % g.gx) * kTx;
  const int i = x0 + lane;
  const int j0 = ((t / g.gx) * kTy + threadIdx.y) * kRows;
  const bool in_x_band = i < g.npml || i >= nx - g.npml;
  const bool fast = x0 + kTx <= nx && j0 >= 1 && j0 + kRows <= ny - 1;
  if (!fast) {
    if (i >= nx) return;
    for (int r = 0; r < kRows; ++r) {
      const int j = j0 + r;
      if (j >= ny) break;
      float ezn, hxn, hyn;
      cell_general<MODE>(g, ez, hx, hy, i, j, ezn, hxn, hyn);
      finish_cell<MODE>(g, i, j, in_x_band, ezn, hxn, hyn, ez_out, hx_out, hy_out);
    }
    return;
  }
  const int a = remap<MODE>(i, nx);
  const int ap = kPer ? (a + 1 == nx ? 0 : a + 1) : min(a + 1, nx - 1);
  const int am = kPer ? (a == 0 ? nx - 1 : a - 1) : max(a - 1, 0);
  // the adjacent lane holds column ap / am exactly when its own column is it
  const bool xp_lane = lane < kTx - 1 && remap<MODE>(i + 1, nx) == ap;
  const bool xm_lane = lane > 0 && remap<MODE>(i - 1, nx) == am;
  const bool upd_hy = kPer || a <= nx - 2;
  const bool upd_e = kPer || a >= 1;
  int b = j0;
  // every row's loads are issued before the march: kRows rows x 3 arrays in
  // flight per wavefront (the march itself only needs the adjacent lanes)
  float ezr[kRows], hxr[kRows], hyr[kRows];
  bool cr[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    ezr[r] = ez[(b + r + 1) * nx + a];
    hxr[r] = hx[(b + r) * nx + a];
    hyr[r] = hy[(b + r) * nx + a];
  }
  if (g.cond != nullptr) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) cr[r] = g.cond[(b + r) * nx + i] != 0;
  } else {
#pragma unroll
    for (int r = 0; r < kRows; ++r) cr[r] = false;
  }
  float e_ym = ez[(b - 1) * nx + a];
  float e_c = ez[b * nx + a];
  // the new Hx at (a, b-1); b-1 <= ny-2 always holds here
  float hxn_m = hx[(b - 1) * nx + a] - g.ch * (e_c - e_ym);
#pragma unroll
  for (int r = 0; r < kRows; ++r, ++b) {
    const float e_yp = ezr[r];
    const float hy0 = hyr[r];
    const float hx0 = hxr[r];
    const float sh_xp = __shfl_down(e_c, 1);
    const float sh_xm = __shfl_up(e_c, 1);
    const float sh_hym = __shfl_up(hy0, 1);
    const float e_xp = xp_lane ? sh_xp : ez[b * nx + ap];
    const float e_xm = xm_lane ? sh_xm : ez[b * nx + am];
    const float hy_m = xm_lane ? sh_hym : hy[b * nx + am];
    const float hyn = upd_hy ? hy0 + g.ch * (e_xp - e_c) : hy0;
    const float hxn = hx0 - g.ch * (e_yp - e_c);
    float ezn = e_c;
    if (upd_e) {
      const float hyn_m = hy_m + g.ch * (e_c - e_xm);
      ezn = e_c + g.ce * ((hyn - hyn_m) - (hxn - hxn_m));
    }
    finish_cell_c<MODE>(g, i, b, in_x_band, cr[r], ezn, hxn, hyn, ez_out, hx_out, hy_out);
    e_ym = e_c;
    e_c = e_yp;
    hxn_m = hxn;
  }
}

static int validate(const Tm2dDesc* d) {
  if (!d) return fail(TM2D_EINVAL, "null descriptor");
  if (d->nx < 3 || d->ny < 3) return fail(TM2D_EINVAL, "grid must be at least 3x3");
  if ((long long)d->nx * d->ny >= (1LL << 31))
    return fail(TM2D_EINVAL, "grid must hold fewer than 2^31 cells");
  if (d->boundary < TM2D_ABSORBING || d->boundary > TM2D_REFLECTING)
    return fail(TM2D_EINVAL, "unknown boundary");
  if (d->boundary == TM2D_ABSORBING && d->npml > 0 && !d->damp)
    return fail(TM2D_EINVAL, "absorbing boundary with npml > 0 needs the damping table");
  if (d->npml < 0) return fail(TM2D_EINVAL, "npml must be >= 0");
  if (d->n_src < 0) return fail(TM2D_EINVAL, "n_src must be >= 0");
  if (d->n_src > 0 && (!d->src_rows || !d->src_table || d->table_steps <= 0))
    return fail(TM2D_EINVAL, "sources need src_rows and src_table");
  return TM2D_OK;
}

// run length of the tile order: an eighth of a tile row (0, the plain
// order, for grids under 8 tiles wide); TM2D_TILE_RUN overrides it
static int tile_run(int gx) {
  if (const char* e = std::getenv("TM2D_TILE_RUN")) {
    const int v = std::atoi(e);
    if (v >= 0 && v <= 4096) return v;
  }
  return gx / kXcds;
}

static int launch(const Tm2dDesc* d, float* const* in, float* const* out, int n,
                  hipStream_t stream) {
  StepArgs g;
  g.nx = d->nx;
  g.ny = d->ny;
  g.npml = d->boundary == TM2D_ABSORBING ? d->npml : 0;
  g.ch = d->ch;
  g.ce = d->ce;
  g.damp = d->damp;
  g.cond = d->conductor;
  g.n_src = d->n_src;
  g.rows = d->src_rows;
  g.vals = d->n_src > 0 ? d->src_table + (long long)n * d->n_src : nullptr;
  const dim3 block(kTx, kTy);
  g.gx = (d->nx + kTx - 1) / kTx;
  const int gy = (d->ny + kTy * kRows - 1) / (kTy * kRows);
  g.ntiles = g.gx * gy;
  g.run = tile_run(g.gx);
  const int grp = g.run > 0 ? kXcds * g.run : 1;
  const dim3 grid((unsigned)((g.ntiles + grp - 1) / grp * grp));
  switch (d->boundary) {
    case TM2D_ABSORBING:
      hipLaunchKernelGGL(tm2d_step_kernel<TM2D_ABSORBING>, grid, block, 0, stream, g, in[0],
                         in[1], in[2], out[0], out[1], out[2]);
      break;
    case TM2D_PERIODIC:
      hipLaunchKernelGGL(tm2d_step_kernel<TM2D_PERIODIC>, grid, block, 0, stream, g, in[0],
                         in[1], in[2], out[0], out[1], out[2]);
      break;
    default:
      hipLaunchKernelGGL(tm2d_step_kernel<TM2D_REFLECTING>, grid, block, 0, stream, g, in[0],
                         in[1], in[2], out[0], out[1], out[2]);
      break;
  }
  TM2D_CHECK(hipGetLastError());
  return TM2D_OK;
}

static int check_fields(float* const* f, const char* what) {
  if (!f || !f[0] || !f[1] || !f[2]) return fail(TM2D_EINVAL, std::string("null ") + what);
  return TM2D_OK;
}

}  // namespace tm2d

using namespace tm2d;

extern "C" {

const char* tm2d_last_error(void) { return g_err.c_str(); }

int tm2d_step(const Tm2dDesc* d, float* const* in, float* const* out, int32_t n,
              void* stream) {
  int rc = validate(d);
  if (rc) return rc;
  if ((rc = check_fields(in, "input fields")) || (rc = check_fields(out, "output fields")))
    return rc;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q)
      if (in[p] == out[q]) return fail(TM2D_EINVAL, "input and output buffers must differ");
  if (d->n_src > 0 && (n < 0 || n >= d->table_steps))
    return fail(TM2D_EINVAL, "step index outside the source table");
  return launch(d, in, out, n, (hipStream_t)stream);
}

int tm2d_run(const Tm2dDesc* d, float* const* a, float* const* b, int32_t n0, int32_t nsteps,
             void* stream) {
  int rc = validate(d);
  if (rc) return rc;
  if ((rc = check_fields(a, "buffer set a")) || (rc = check_fields(b, "buffer set b"))) return rc;
  for (int p = 0; p < 3; ++p)
    for (int q = 0; q < 3; ++q)
      if (a[p] == b[q]) return fail(TM2D_EINVAL, "the two buffer sets must differ");
  if (nsteps < 0 || n0 < 0) return fail(TM2D_EINVAL, "negative step range");
  if (d->n_src > 0 && (long long)n0 + nsteps > d->table_steps)
    return fail(TM2D_EINVAL, "step range outside the source table");
  for (int s = 0; s < nsteps; ++s) {
    float* const* src = (s % 2 == 0) ? a : b;
    float* const* dst = (s % 2 == 0) ? b : a;
    if ((rc = launch(d, src, dst, n0 + s, (hipStream_t)stream))) return rc;
  }
  return TM2D_OK;
}

}  // extern "C"
