// Internal definitions shared by the forward (yee.hip) and adjoint
// (yee_adjoint.hip) translation units of libyee_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/yee.h"

namespace yee {

// thread-local message behind yee_last_error() (defined in yee.hip)
extern thread_local std::string g_err;

inline int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define YEE_HIP_CHECK(expr)                                                   \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess)                                                     \
      return fail(YEE_EHIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int kBX = 64;   // x extent of a block: one wavefront
constexpr int kBY = 4;    // y rows per block
constexpr int kThreads = kBX * kBY;

enum Mode { kWalls = 0, kSponge = 1, kPeriodic = 2, kCpml = 3 };

// psi array slots (CPML).  y-slab arrays are (nz, 2p, nx), z-slab arrays
// (2p, ny, nx), x-slab arrays (nz, ny, 2p), all x fastest.
enum Psi {
  kExy = 0, kEzy, kHxy, kHzy,   // derivative along y
  kExz, kEyz, kHxz, kHyz,       // along z
  kEyx, kEzx, kHyx, kHzx,       // along x
  kNumPsi
};

struct GridArgs {
  int nx, ny, nz, p;
  float ch;
  float* f[6];
  const float* cb;
  const float* ca;
  const float* ssx;
  const float* ssy;
  const float* ssz;
  float sdt, seps0;
  // CPML tables per axis: bE cE kE bH cH kH, each n_a long
  const float* cpx;
  const float* cpy;
  const float* cpz;
  float* psi[kNumPsi];
  const unsigned char* mat;  // material palette (see YeeGridDesc)
  const float* pal;
};

struct SrcArgs {
  int nsrc;  // sources handled inside the E kernel (<= YEE_FUSED_SOURCES)
  int comp[YEE_FUSED_SOURCES];
  long long idx[YEE_FUSED_SOURCES];
  float val[YEE_FUSED_SOURCES];
};

// What an E launch records of the step it computes (the forward run of the
// adjoint): on a box, d = E^{n+1} - Ca E^n (comp-major, 3 * bc floats, x
// fastest; the difference the eps gradient needs), followed, for lossy
// plans, by E^n (3 * bc more); and the values of probes given by a cell ->
// probe map over their bounding box,
// pmap[((c * pnz + k) * pny + j) * pnx + i] = probe of E component c or -1.
struct RecArgs {
  float* box_out;  // null: no box
  int bx0, by0, bz0, bnx, bny, bnz;
  unsigned bc;
  float* probe_out;  // null: no probes
  const int* pmap;
  int px0, py0, pz0, pnx, pny, pnz;
};

inline __device__ int slab_index(int a, int n, int p) {
  return a < p ? a : a - (n - 2 * p);
}

// Tiles of a z-marching launch: gx * gy * gz tiles of 64 x 4 columns by kz
// planes, numbered x fastest, then y, then z chunk.  The launch is 1-D.
// Workgroups are dispatched round-robin over the chip's 8 XCDs (workgroup L
// runs on XCD L % 8), each XCD with its own L2.  With `per` > 0 the
// workgroup -> tile map is t = (L % 8) * per + L / 8: every XCD works
// through its own contiguous range of tiles in order, so the x/y
// neighbour tiles whose edge rows and columns a tile re-reads sit in the
// same L2 at about the same time.  per == 0 is the plain order t = L.
constexpr int kXcds = 8;

struct Tiling {
  int gx, gy, ntiles, per, run;
};

// order: 0 plain, 1 (or -1) XCD-contiguous, >= 3 XCD row runs of
// (order - 2) rows of tiles (see tileset_finish)
inline Tiling make_tiling(int nx, int ny, int nz, int kz, int order) {
  Tiling t;
  t.gx = (nx + 63) / 64;
  t.gy = (ny + 3) / 4;
  const int gz = (nz + kz - 1) / kz;
  t.ntiles = t.gx * t.gy * gz;
  t.run = order >= 3 ? (order - 2) * t.gx : 0;
  t.per = (order != 0 && t.run == 0) ? (t.ntiles + kXcds - 1) / kXcds : 0;
  return t;
}

inline unsigned tiling_blocks(const Tiling& t) {
  if (t.run > 0) {
    const int g = kXcds * t.run;
    return (unsigned)((t.ntiles + g - 1) / g * g);
  }
  return t.per > 0 ? (unsigned)(t.per * kXcds) : (unsigned)t.ntiles;
}


// A set of tiles given as up to kMaxBoxes boxes of cells [x0, x1) x
// [y0, y1) x [z0, z1), each cut into tiles of 64 columns x `rows` rows x kz
// planes (edge tiles partial); `first` is the set-wide index of a box's
// first tile.  Used to launch the CPML interior and its slab shell
// separately.
constexpr int kMaxBoxes = 6;

struct TileBox {
  int x0, x1, nxt, y0, y1, nyt, z0, z1, first;
  // paired box (xh1 > xh0): one 64-lane tile column holds [x0, x1) in lanes
  // 0-31 and [xh0, xh1) in lanes 32-63 (both at most 32 wide)
  int xh0, xh1;
};

struct TileSet {
  int nbox, ntiles, per, kz, rows;
  int band, bq;  // banded XCD order: tiles per z-chunk band, per-XCD share
  int run;       // XCD row runs: tiles per run (0: not this order)
  TileBox b[kMaxBoxes];
};

inline void tileset_add(TileSet& ts, int x0, int x1, int y0, int y1, int z0, int z1,
                        int xh0 = 0, int xh1 = 0) {
  if (x1 <= x0 || y1 <= y0 || z1 <= z0) return;
  TileBox& B = ts.b[ts.nbox++];
  B.x0 = x0;
  B.x1 = x1;
  B.xh0 = xh0;
  B.xh1 = xh1;
  B.nxt = xh1 > xh0 ? 1 : (x1 - x0 + 63) / 64;
  B.y0 = y0;
  B.y1 = y1;
  B.nyt = (y1 - y0 + ts.rows - 1) / ts.rows;
  B.z0 = z0;
  B.z1 = z1;
  B.first = ts.ntiles;
  ts.ntiles += B.nxt * B.nyt * ((z1 - z0 + ts.kz - 1) / ts.kz);
}

// Tile order of a set: 0 plain (t = L), 1 XCD-contiguous (as Tiling), 2
// XCD-banded (single-box sets): the tiles of one z-chunk band are split into
// 8 contiguous shares, one per XCD, and every XCD moves through the bands
// at the same pace, so all XCDs stream the same planes while each keeps
// its neighbouring tiles in its own L2.
//
// Orders 3 and up: XCD row runs.  Consecutive groups of 8 * run workgroups
// cover 8 * run consecutive tiles; within a group XCD x takes the run of
// `run` tiles starting at x * run, with run = (order - 2) rows of tiles of
// the first box.  Each XCD then computes whole rows of x-neighbouring tiles
// (their shared edge columns stay in its L2) while the 8 XCDs stream
// adjacent runs of memory at the same time.
inline void tileset_finish(TileSet& ts, int order) {
  ts.per = 0;
  ts.band = ts.bq = 0;
  ts.run = 0;
  if (order >= 3 && ts.ntiles > 0) {
    ts.run = (order - 2) * (ts.b[0].nxt > 0 ? ts.b[0].nxt : 1);
  } else if (order == 2 && ts.nbox == 1 && ts.ntiles > 0) {
    ts.band = ts.b[0].nxt * ts.b[0].nyt;
    ts.bq = (ts.band + kXcds - 1) / kXcds;
  } else if (order >= 1) {
    ts.per = (ts.ntiles + kXcds - 1) / kXcds;
  }
}

inline unsigned tileset_blocks(const TileSet& ts) {
  if (ts.run > 0) {
    const int g = kXcds * ts.run;
    return (unsigned)((ts.ntiles + g - 1) / g * g);
  }
  if (ts.band > 0) {
    const int nb = (ts.ntiles + ts.band - 1) / ts.band;
    return (unsigned)(nb * ts.bq * kXcds);
  }
  return ts.per > 0 ? (unsigned)(ts.per * kXcds) : (unsigned)ts.ntiles;
}

// Split a grid into `in`, the cells whose neighbourhood -- `lo` cells
// towards lower and `hi` towards higher x and y indices, zlo / zhi along
// z -- meets no CPML slab of thickness p, as one box whose x range is
// aligned to 32 columns (128-byte rows); and `sh`, every other cell, as up
// to six boxes.  p == 0 puts the whole grid in `in`.
inline void split_tiles(int nx, int ny, int nz, int p, int kz, int lo, int hi, int zlo, int zhi,
                        int order, TileSet& in, TileSet& sh, int rows = 4, bool pair_x = false) {
  std::memset(&in, 0, sizeof(in));
  std::memset(&sh, 0, sizeof(sh));
  in.kz = sh.kz = kz;
  in.rows = sh.rows = rows;
  if (p <= 0) {
    tileset_add(in, 0, nx, 0, ny, 0, nz);
  } else {
    const int xa = (p + lo + 31) / 32 * 32, xb = (nx - p - hi) / 32 * 32;
    const int ya = p + lo, yb = ny - p - hi;
    const int za = p + zlo, zb = nz - p - zhi;
    if (xa < xb && ya < yb && za < zb) {
      tileset_add(in, xa, xb, ya, yb, za, zb);
      tileset_add(sh, 0, nx, 0, ny, 0, za);
      tileset_add(sh, 0, nx, 0, ny, zb, nz);
      tileset_add(sh, 0, nx, 0, ya, za, zb);
      tileset_add(sh, 0, nx, yb, ny, za, zb);
      if (pair_x && xa <= 32 && nx - xb <= 32) {
        tileset_add(sh, 0, xa, ya, yb, za, zb, xb, nx);
      } else {
        tileset_add(sh, 0, xa, ya, yb, za, zb);
        tileset_add(sh, xb, nx, ya, yb, za, zb);
      }
    } else {
      tileset_add(sh, 0, nx, 0, ny, 0, nz);
    }
  }
  tileset_finish(in, order);
  tileset_finish(sh, order == 2 ? 0 : (order > 2 ? 3 : order));
}

// Three-way split for the fused step: `in` as in split_tiles (x aligned to
// 32 columns), `shx` the x-slab columns [0, xa) and [xb, nx) over all y and
// z (one paired box when both are at most 32 wide), and `shn` the rest of
// the shell -- y-slab rows and z-slab planes over x in [xa, xb), where no
// cell or x-1 neighbour lies in an x slab.  When there is no interior the
// whole grid goes to `shx`.
inline void split_tiles3(int nx, int ny, int nz, int p, int kz, int lo, int hi, int order,
                         TileSet& in, TileSet& shn, TileSet& shx, int rows, int zlo = 0,
                         int zhi = 0, bool pair_x = true, int shell_order = -1) {
  std::memset(&in, 0, sizeof(in));
  std::memset(&shn, 0, sizeof(shn));
  std::memset(&shx, 0, sizeof(shx));
  in.kz = shn.kz = shx.kz = kz;
  in.rows = shn.rows = shx.rows = rows;
  const int xa = (p + lo + 31) / 32 * 32, xb = (nx - p - hi) / 32 * 32;
  const int ya = p + lo, yb = ny - p - hi;
  const int za = p + zlo, zb = nz - p - zhi;
  if (p <= 0) {
    tileset_add(in, 0, nx, 0, ny, 0, nz);
  } else if (xa < xb && ya < yb && za < zb) {
    tileset_add(in, xa, xb, ya, yb, za, zb);
    if (pair_x && xa <= 32 && nx - xb <= 32) {
      tileset_add(shx, 0, xa, 0, ny, 0, nz, xb, nx);
    } else {
      tileset_add(shx, 0, xa, 0, ny, 0, nz);
      tileset_add(shx, xb, nx, 0, ny, 0, nz);
    }
    tileset_add(shn, xa, xb, 0, ya, 0, nz);
    tileset_add(shn, xa, xb, yb, ny, 0, nz);
    tileset_add(shn, xa, xb, ya, yb, 0, za);
    tileset_add(shn, xa, xb, ya, yb, zb, nz);
  } else {
    tileset_add(shx, 0, nx, 0, ny, 0, nz);
  }
  tileset_finish(in, order);
  const int so = shell_order >= 0 ? shell_order : (order == 2 ? 0 : (order > 2 ? 3 : order));
  tileset_finish(shn, so);
  tileset_finish(shx, so);
}

// Position of a workgroup's tile: first column / row / plane, the exclusive
// limits of its box (x, y) and of the tile (k1), and for a paired box the
// high half (lanes 32-63: columns ih0 .. ihl-1).
struct TilePos {
  int i0, xl, j0, yl, k0, k1, ih0, ihl;
  bool pair;
};

// false for the padding workgroups of the XCD-aware order (same map as
// Tiling)
__device__ __forceinline__ bool tile_of_set_at(const TileSet& ts, int L, TilePos& tp) {
  int t;
  if (ts.run > 0) {
    const int g = kXcds * ts.run;
    t = (L / g) * g + (L % kXcds) * ts.run + (L / kXcds) % ts.run;
  } else if (ts.band > 0) {
    const int s = L / kXcds;
    const int loc = (L % kXcds) * ts.bq + s % ts.bq;
    if (loc >= ts.band) return false;
    t = (s / ts.bq) * ts.band + loc;
  } else {
    t = ts.per > 0 ? (L % kXcds) * ts.per + L / kXcds : L;
  }
  if (t >= ts.ntiles) return false;
  TileBox B = ts.b[0];
#pragma unroll
  for (int q = 1; q < kMaxBoxes; ++q)
    if (q < ts.nbox && t >= ts.b[q].first) B = ts.b[q];
  const int lt = t - B.first;
  tp.i0 = B.x0 + (lt % B.nxt) * 64;
  tp.xl = B.x1;
  tp.pair = B.xh1 > B.xh0;
  tp.ih0 = B.xh0;
  tp.ihl = B.xh1;
  const int r = lt / B.nxt;
  tp.j0 = B.y0 + (r % B.nyt) * ts.rows;
  tp.yl = B.y1;
  tp.k0 = B.z0 + (r / B.nyt) * ts.kz;
  tp.k1 = min(tp.k0 + ts.kz, B.z1);
  return true;
}

__device__ __forceinline__ bool tile_of_set(const TileSet& ts, TilePos& tp) {
  return tile_of_set_at(ts, (int)blockIdx.x, tp);
}

// (bx, by, bz) of this workgroup; false for the padding workgroups of the
// remapped order
__device__ __forceinline__ bool tile_of(const Tiling& tl, int& bx, int& by, int& bz) {
  const int L = blockIdx.x;
  int t;
  if (tl.run > 0) {
    const int g = kXcds * tl.run;
    t = (L / g) * g + (L % kXcds) * tl.run + (L / kXcds) % tl.run;
  } else {
    t = tl.per > 0 ? (L % kXcds) * tl.per + L / kXcds : L;
  }
  if (t >= tl.ntiles) return false;
  bx = t % tl.gx;
  const int r = t / tl.gx;
  by = r % tl.gy;
  bz = r / tl.gy;
  return true;
}

}  // namespace yee

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct YeePlan {
  YeeGridDesc d;
  int kz;            // z planes per block of the two-pass kernels
  int adj_kz = 6;    // z planes per block of the adjoint kernels (160^3
                     // design workload, software-pipelined bodies with
                     // preloaded psi: 2, 3, 4, 6, 8 planes give 3.42,
                     // 3.25, 3.22, 3.19, 3.22x one forward run in the
                     // checkpointed mode)
  int fused_kz = 0;  // z planes per block of the fused kernel (0: automatic)
  int adj_pipe = 1;  // adjoint kernels: software-pipelined bodies (1) or
                     // loads issued per plane (0); same results
  int fused_variant = 3;  // 0: LDS + halo row, 1-3: barrier-free (see yee_fused.hip)
  int fused_rows = 4;     // y rows (wavefronts) per tile of variants 2 and 3: 4, 8 or 16
  int xcd = -1;      // tile order: 0 plain, 1 XCD-contiguous, 2 XCD-banded, -1 automatic
  int xcd_shell = -1;  // tile order of the fused CPML shell launches (-1: from xcd)
  // sources
  int nsrc = 0;
  int nsteps_table = 0;
  std::vector<int> comp;
  std::vector<long long> idx;
  std::vector<float> table;  // host copy, step-major
  int* d_comp = nullptr;     // device copies for the unfused tail
  long long* d_idx = nullptr;
  float* d_table = nullptr;
  // device copy of the adjoint's CPML tile sets and the host bytes it holds
  // (a cache of what the launches derive from the fields above)
  mutable void* d_adjsets = nullptr;
  mutable std::vector<unsigned char> h_adjsets;
};


namespace yee {

inline long long psi_slab_floats(const YeeGridDesc& d, int which /*0 y,1 z,2 x*/) {
  const long long tp = 2LL * d.npml;
  if (which == 0) return (long long)d.nz * tp * d.nx;
  if (which == 1) return tp * d.ny * d.nx;
  return (long long)d.nz * d.ny * tp;
}

inline GridArgs make_args(const YeePlan* p, float* const* fields, float* psi) {
  const YeeGridDesc& d = p->d;
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  g.nx = d.nx;
  g.ny = d.ny;
  g.nz = d.nz;
  g.p = d.boundary == YEE_BOUNDARY_CPML ? d.npml : 0;
  g.ch = d.ch;
  for (int c = 0; c < 6; ++c) g.f[c] = fields[c];
  g.cb = d.cb;
  g.ca = d.ca;
  g.mat = d.mat;
  g.pal = d.mat ? d.palette : nullptr;
  if (d.boundary == YEE_BOUNDARY_SPONGE) {
    g.ssx = d.sponge_sigma;
    g.ssy = d.sponge_sigma + d.nx;
    g.ssz = d.sponge_sigma + d.nx + d.ny;
    g.sdt = d.sponge_dt;
    g.seps0 = d.sponge_eps0;
  }
  if (d.boundary == YEE_BOUNDARY_CPML) {
    g.cpx = d.cpml;
    g.cpy = d.cpml + 6LL * d.nx;
    g.cpz = d.cpml + 6LL * (d.nx + d.ny);
    const long long ny_ = psi_slab_floats(d, 0), nz_ = psi_slab_floats(d, 1),
                    nx_ = psi_slab_floats(d, 2);
    float* q = psi;
    for (int a = kExy; a <= kHzy; ++a) { g.psi[a] = q; q += ny_; }
    for (int a = kExz; a <= kHyz; ++a) { g.psi[a] = q; q += nz_; }
    for (int a = kEyx; a <= kHzx; ++a) { g.psi[a] = q; q += nx_; }
  }
  return g;
}

inline int check_fields(const YeePlan* p, float* const* fields, float* psi) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (!fields) return fail(YEE_EINVAL, "null field array");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(YEE_EINVAL, "null field pointer " + std::to_string(c));
  if (p->d.boundary == YEE_BOUNDARY_CPML && !psi)
    return fail(YEE_EINVAL, "cpml needs the psi buffer");
  return YEE_OK;
}

// sources of step n fused into an E update (the first YEE_FUSED_SOURCES)
inline int make_src_args(const YeePlan* p, int n, SrcArgs& s) {
  std::memset(&s, 0, sizeof(s));
  if (p->nsrc > 0 && (n < 0 || n >= p->nsteps_table))
    return fail(YEE_EINVAL, "step " + std::to_string(n) + " outside the source table (" +
                                std::to_string(p->nsteps_table) + " steps)");
  s.nsrc = p->nsrc < YEE_FUSED_SOURCES ? p->nsrc : YEE_FUSED_SOURCES;
  for (int q = 0; q < s.nsrc; ++q) {
    s.comp[q] = p->comp[q];
    s.idx[q] = p->idx[q];
    s.val[q] = p->table[(size_t)n * p->nsrc + q];
  }
  return YEE_OK;
}

// forward launches (yee.hip)
// sources first..nsrc-1 of step n, one thread, insertion order
int launch_inject(const YeePlan* p, float* ex, float* ey, float* ez, int n, int first,
                  hipStream_t st);
int launch_h(const YeePlan* p, const GridArgs& g, hipStream_t st);
// rec (optional): record inside the E kernel; only when every source of the
// plan is fused into it (no injection launch after the kernel)
int launch_e(const YeePlan* p, const GridArgs& g, int n, hipStream_t st,
             const RecArgs* rec = nullptr);

}  // namespace yee
