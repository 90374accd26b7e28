// jaxphys_amd: fused single-pass Yee step for gfx950 -- both half-steps of
// src/neurosim/em/fdtd3d.py:280-379 in ONE launch per time step.
//
// The two-pass kernels of yee.hip stream E and H twice per step (76 B per
// cell-update, 80 B with Ca).  Here every field is read once from buffer A
// and written once to buffer B (52 / 56 B): E^n, H^{n-1/2}, psi come from
// A; H^{n+1/2}, E^{n+1}, psi' go to B (ping-pong, so a block never reads a
// value another block has already overwritten).
//
// Block = (BY+1) wavefronts x 64 lanes; lane = x, wave = y row.  Wave 0 is a
// halo row: it computes H^{n+1/2} on row j0-1 only, so that the owner rows
// j0..j0+BY-1 find H(y-1) in LDS.  Lane 0 of every owner wave recomputes
// Hy, Hz of column x-1 (the previous tile's last column).  H(z-1) is the
// thread's own H of the previous plane (register carry); each block marches
// a chunk of z planes after recomputing Hx, Hy of the plane below the chunk.
// Per plane: own H -> LDS (double-buffered by plane parity) and to B; one
// barrier; owners compute E^{n+1} from LDS + registers -> B.  Point sources
// are added by a one-thread kernel after the step (insertion order, after
// the E update, as the reference does).  Arithmetic and its order are those
// of the two-pass kernels, so the result is bit-for-bit the same.
//
// Kernel arguments are kept compact (one base pointer for each psi buffer,
// one for all CPML tables, one for the sponge profiles) and offsets are
// 32-bit (plans are limited to < 2^32 cells): the address arithmetic then
// needs one vector register per neighbour instead of two.

#include "yee_internal.h"

namespace yee {

struct FusedArgs {
  int nx, ny, nz, p;
  float ch, sdt, seps0;
  long long ysl, zsl, xsl;      // psi slab sizes (floats) of the y, z, x groups
  const float* fi[6];           // buffer A fields
  float* fo[6];                 // buffer B fields
  const float* cb;
  const float* ca;
  const float* sponge;          // sigma_x | sigma_y | sigma_z
  const float* cpml;            // x | y | z tables, 6 n_a floats each
  const float* pin;             // psi of buffer A
  float* pout;                  // psi of buffer B
  const unsigned char* mat;     // material palette (variants 3-4; see YeeGridDesc)
  const float* pal;
};

// psi slot offsets: y group (Exy Ezy Hxy Hzy), z group (Exz Eyz Hxz Hyz),
// x group (Eyx Ezx Hyx Hzx) -- the slot order of yee_internal.h
__device__ __forceinline__ long long slot_base(const FusedArgs& a, int s) {
  if (s < 4) return s * a.ysl;
  if (s < 8) return 4 * a.ysl + (s - 4) * a.zsl;
  return 4 * a.ysl + 4 * a.zsl + (s - 8) * a.xsl;
}

__device__ __forceinline__ const float* tab_of(const FusedArgs& a, int ax) {
  return ax == 0 ? a.cpml : (ax == 1 ? a.cpml + 6LL * a.nx : a.cpml + 6LL * (a.nx + a.ny));
}

// Stores of a step's outputs (fields and psi) are non-temporal: each value
// is written once and read again only by the next step, after the whole
// grid has streamed through the caches, so keeping it there only evicts
// the input lines that neighbouring tiles still read.  Measured on the
// 512^3 CPML workload: 1.911 -> 1.779 ms per fused step (interior launch
// 1134 -> 1049 us), the same bits.
__device__ __forceinline__ void put(float* p, float v) { __builtin_nontemporal_store(v, p); }

// psi' = b psi + c d (psi from A, written to B when `write`); returns d/kappa + psi'
__device__ __forceinline__ float cterm(const FusedArgs& a, int slot, unsigned off, bool write,
                                       int ax, int n, int v, int kind, float d) {
  const float* t = tab_of(a, ax);
  const float b = t[(3 * kind + 0) * n + v];
  const float c = t[(3 * kind + 1) * n + v];
  const float kinv = t[(3 * kind + 2) * n + v];
  const unsigned o = (unsigned)slot_base(a, slot) + off;
  const float ps = b * a.pin[o] + c * d;
  if (write) put(&a.pout[o], ps);
  return d * kinv + ps;
}

// H^{n+1/2} at (i, j, k) from buffer A.  WANT bit 1/2/4: Hx/Hy/Hz.  Writes
// psi_H' to B when `own`.  ``e`` receives Ex, Ey, Ez of the cell.
template <int MODE, int WANT>
__device__ __forceinline__ void h_cell(const FusedArgs& a, int i, int j, int k, bool own,
                                       float& ox, float& oy, float& oz, float* e) {
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  constexpr bool per = MODE == kPeriodic;
  const unsigned P = (unsigned)nx * ny;
  const unsigned c = ((unsigned)k * ny + j) * nx + i;
  const bool xin = per || i < nx - 1;
  const bool yin = per || j < ny - 1;
  const bool zin = per || k < nz - 1;
  const unsigned di = (per && i == nx - 1) ? -(unsigned)(nx - 1) : 1;
  const unsigned dj = (per && j == ny - 1) ? -(unsigned)(ny - 1) * nx : nx;
  const unsigned dk = (per && k == nz - 1) ? -(unsigned)(nz - 1) * P : P;
  const float* __restrict__ ex = a.fi[0];
  const float* __restrict__ ey = a.fi[1];
  const float* __restrict__ ez = a.fi[2];
  const float ex0 = ex[c], ey0 = ey[c], ez0 = ez[c];
  if (e) {
    e[0] = ex0;
    e[1] = ey0;
    e[2] = ez0;
  }
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  const bool zs = MODE == kCpml && (k < p || k >= nz - p);
  const unsigned oy_ = ys ? ((unsigned)(k * 2 * p + slab_index(j, ny, p)) * nx + i) : 0;
  const unsigned oz_ = zs ? ((unsigned)(slab_index(k, nz, p) * ny + j) * nx + i) : 0;
  const unsigned ox_ = xs ? (((unsigned)k * ny + j) * 2 * p + slab_index(i, nx, p)) : 0;
  const float ch = a.ch;
  // branch-free: out-of-range neighbours are read at the cell itself and the
  // update is then not selected
  const unsigned dic = xin ? di : 0u, djc = yin ? dj : 0u, dkc = zin ? dk : 0u;
  if (WANT & 1) {
    const float hx = a.fi[3][c];
    float u = ez[c + djc] - ez0, v = ey[c + dkc] - ey0;  // fdtd3d.py:305-312
    if (MODE == kCpml && yin && zin) {
      if (ys) u = cterm(a, kHxy, oy_, own, 1, ny, j, 1, u);
      if (zs) v = cterm(a, kHxz, oz_, own, 2, nz, k, 1, v);
    }
    ox = (yin && zin) ? hx - ch * (u - v) : hx;
  }
  if (WANT & 2) {
    const float hy = a.fi[4][c];
    float u = ex[c + dkc] - ex0, v = ez[c + dic] - ez0;  // fdtd3d.py:314-321
    if (MODE == kCpml && xin && zin) {
      if (zs) u = cterm(a, kHyz, oz_, own, 2, nz, k, 1, u);
      if (xs) v = cterm(a, kHyx, ox_, own, 0, nx, i, 1, v);
    }
    oy = (xin && zin) ? hy - ch * (u - v) : hy;
  }
  if (WANT & 4) {
    const float hz = a.fi[5][c];
    float u = ey[c + dic] - ey0, v = ex[c + djc] - ex0;  // fdtd3d.py:323-330
    if (MODE == kCpml && xin && yin) {
      if (xs) u = cterm(a, kHzx, ox_, own, 0, nx, i, 1, u);
      if (ys) v = cterm(a, kHzy, oy_, own, 1, ny, j, 1, v);
    }
    oz = (xin && yin) ? hz - ch * (u - v) : hz;
  }
}

template <int MODE, bool LOSSY, int BY>
__global__ __launch_bounds__(64 * (BY + 1)) void yee_fused_kernel(FusedArgs a, int kz) {
  __shared__ float shx[2][BY + 1][64];
  __shared__ float shy[2][BY + 1][64];
  __shared__ float shz[2][BY + 1][64];
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  constexpr bool per = MODE == kPeriodic;
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const int i = blockIdx.x * 64 + lane;
  const int jr = blockIdx.y * BY + w - 1;
  int j = jr;
  bool jvalid;
  if (per) {
    if (j < 0) j += ny;
    jvalid = jr < ny;
  } else {
    jvalid = j >= 0 && j < ny;
  }
  const bool valid = i < nx && jvalid;
  const bool owner = valid && w >= 1;
  int im = i - 1;
  bool imvalid = im >= 0;
  if (per && im < 0) {
    im += nx;
    imvalid = true;
  }
  const int k0 = blockIdx.z * kz;
  const int k1 = min(k0 + kz, nz);
  const unsigned P = (unsigned)nx * ny;
  // Hx, Hy of the plane below the chunk (the z-1 neighbour of the first plane)
  float hxm = 0.f, hym = 0.f;
  {
    int kp = k0 - 1;
    if (per && kp < 0) kp += nz;
    if (owner && kp >= 0) {
      float unused;
      h_cell<MODE, 3>(a, i, j, kp, false, hxm, hym, unused, nullptr);
    }
  }
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  float sxy = 0.f;
  if (MODE == kSponge && owner) sxy = a.sponge[i] + a.sponge[nx + j];
  for (int k = k0; k < k1; ++k) {
    const int buf = k & 1;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    float e[3] = {0.f, 0.f, 0.f};
    if (valid) h_cell<MODE, 7>(a, i, j, k, owner, hx, hy, hz, e);
    shx[buf][w][lane] = hx;
    shy[buf][w][lane] = hy;
    shz[buf][w][lane] = hz;
    float hyxm = 0.f, hzxm = 0.f;
    if (owner && lane == 0 && imvalid) {
      float unused;
      h_cell<MODE, 6>(a, im, j, k, false, unused, hyxm, hzxm, nullptr);
    }
    const unsigned c = (unsigned)k * P + (unsigned)j * nx + i;
    if (owner) {
      a.fo[3][c] = hx;
      a.fo[4][c] = hy;
      a.fo[5][c] = hz;
    }
    __syncthreads();
    if (owner) {
      const float hzj = shz[buf][w - 1][lane];
      const float hxj = shx[buf][w - 1][lane];
      const float hzi = lane > 0 ? shz[buf][w][lane - 1] : hzxm;
      const float hyi = lane > 0 ? shy[buf][w][lane - 1] : hyxm;
      const bool xin = per || i > 0;
      const bool yin = per || j > 0;
      const bool zin = per || k > 0;
      const float cb = a.cb[c];
      const float ca = LOSSY ? a.ca[c] : 1.f;
      float vx = e[0], vy = e[1], vz = e[2];
      const bool zs = MODE == kCpml && (k < p || k >= nz - p);
      const unsigned oy_ = ys ? ((unsigned)(k * 2 * p + slab_index(j, ny, p)) * nx + i) : 0;
      const unsigned oz_ = zs ? ((unsigned)(slab_index(k, nz, p) * ny + j) * nx + i) : 0;
      const unsigned ox_ = xs ? (((unsigned)k * ny + j) * 2 * p + slab_index(i, nx, p)) : 0;
      if (yin && zin) {  // fdtd3d.py:334-341
        float u = hz - hzj, v = hy - hym;
        if (MODE == kCpml) {
          if (ys) u = cterm(a, kExy, oy_, true, 1, ny, j, 0, u);
          if (zs) v = cterm(a, kExz, oz_, true, 2, nz, k, 0, v);
        }
        vx = LOSSY ? ca * vx + cb * (u - v) : vx + cb * (u - v);
      }
      if (xin && zin) {  // fdtd3d.py:343-350
        float u = hx - hxm, v = hz - hzi;
        if (MODE == kCpml) {
          if (zs) u = cterm(a, kEyz, oz_, true, 2, nz, k, 0, u);
          if (xs) v = cterm(a, kEyx, ox_, true, 0, nx, i, 0, v);
        }
        vy = LOSSY ? ca * vy + cb * (u - v) : vy + cb * (u - v);
      }
      if (xin && yin) {  // fdtd3d.py:352-359
        float u = hy - hyi, v = hx - hxj;
        if (MODE == kCpml) {
          if (xs) u = cterm(a, kEzx, ox_, true, 0, nx, i, 0, u);
          if (ys) v = cterm(a, kEzy, oy_, true, 1, ny, j, 0, v);
        }
        vz = LOSSY ? ca * vz + cb * (u - v) : vz + cb * (u - v);
      }
      if (MODE == kSponge) {  // fdtd3d.py:361-366
        const float sig = (sxy + a.sponge[nx + ny + k]) / 3.0f;
        if (sig != 0.f) {
          const float d = expf(((-sig) * a.sdt) / a.seps0);
          vx *= d;
          vy *= d;
          vz *= d;
        }
      }
      a.fo[0][c] = vx;
      a.fo[1][c] = vy;
      a.fo[2][c] = vz;
    }
    hxm = hx;
    hym = hy;
  }
}

// Barrier-free variant: 64 x 4 threads, every thread owns its column and
// recomputes, besides its own H^{n+1/2}, Hy and Hz of column x-1 and Hx and
// Hz of row y-1 from buffer A (their inputs are the thread's neighbours, so
// the extra loads hit L1/L2).  Straight-line code: every load of a plane is
// in flight at once and no wave ever waits for another.
template <int MODE, bool LOSSY>
__global__ __launch_bounds__(256) void yee_fused_nolds_kernel(
    FusedArgs a, int kz, const float* __restrict__ fi0, const float* __restrict__ fi1,
    const float* __restrict__ fi2, const float* __restrict__ fi3, const float* __restrict__ fi4,
    const float* __restrict__ fi5, float* __restrict__ fo0, float* __restrict__ fo1,
    float* __restrict__ fo2, float* __restrict__ fo3, float* __restrict__ fo4,
    float* __restrict__ fo5) {
  a.fi[0] = fi0; a.fi[1] = fi1; a.fi[2] = fi2; a.fi[3] = fi3; a.fi[4] = fi4; a.fi[5] = fi5;
  a.fo[0] = fo0; a.fo[1] = fo1; a.fo[2] = fo2; a.fo[3] = fo3; a.fo[4] = fo4; a.fo[5] = fo5;
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  constexpr bool per = MODE == kPeriodic;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 64 + lane;
  const int j = blockIdx.y * 4 + (threadIdx.x >> 6);
  if (i >= nx || j >= ny) return;
  int im = i - 1, jm = j - 1;
  bool imvalid = im >= 0, jmvalid = jm >= 0;
  if (per && im < 0) { im += nx; imvalid = true; }
  if (per && jm < 0) { jm += ny; jmvalid = true; }
  // out-of-grid neighbours (non-periodic) are never used by an in-range
  // update; computing them at a clamped index keeps the code branch-free
  const int imc = imvalid ? im : i, jmc = jmvalid ? jm : j;
  const int k0 = blockIdx.z * kz;
  const int k1 = min(k0 + kz, nz);
  const unsigned P = (unsigned)nx * ny;
  float hxm = 0.f, hym = 0.f;
  {
    int kp = k0 - 1;
    if (per && kp < 0) kp += nz;
    if (kp >= 0) {
      float unused;
      h_cell<MODE, 3>(a, i, j, kp, false, hxm, hym, unused, nullptr);
    }
  }
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  float sxy = 0.f;
  if (MODE == kSponge) sxy = a.sponge[i] + a.sponge[nx + j];
  const bool xin = per || i > 0;
  const bool yin = per || j > 0;
  for (int k = k0; k < k1; ++k) {
    float hx, hy, hz, unused, hyi, hzi, hxj, hzj;
    float e[3];
    h_cell<MODE, 7>(a, i, j, k, true, hx, hy, hz, e);
    h_cell<MODE, 6>(a, imc, j, k, false, unused, hyi, hzi, nullptr);
    h_cell<MODE, 5>(a, i, jmc, k, false, hxj, unused, hzj, nullptr);
    const unsigned c = (unsigned)k * P + (unsigned)j * nx + i;
    a.fo[3][c] = hx;
    a.fo[4][c] = hy;
    a.fo[5][c] = hz;
    const bool zin = per || k > 0;
    const float cb = a.cb[c];
    const float ca = LOSSY ? a.ca[c] : 1.f;
    float vx = e[0], vy = e[1], vz = e[2];
    const bool zs = MODE == kCpml && (k < p || k >= nz - p);
    const unsigned oy_ = ys ? ((unsigned)(k * 2 * p + slab_index(j, ny, p)) * nx + i) : 0;
    const unsigned oz_ = zs ? ((unsigned)(slab_index(k, nz, p) * ny + j) * nx + i) : 0;
    const unsigned ox_ = xs ? (((unsigned)k * ny + j) * 2 * p + slab_index(i, nx, p)) : 0;
    if (yin && zin) {
      float u = hz - hzj, v = hy - hym;
      if (MODE == kCpml) {
        if (ys) u = cterm(a, kExy, oy_, true, 1, ny, j, 0, u);
        if (zs) v = cterm(a, kExz, oz_, true, 2, nz, k, 0, v);
      }
      vx = LOSSY ? ca * vx + cb * (u - v) : vx + cb * (u - v);
    }
    if (xin && zin) {
      float u = hx - hxm, v = hz - hzi;
      if (MODE == kCpml) {
        if (zs) u = cterm(a, kEyz, oz_, true, 2, nz, k, 0, u);
        if (xs) v = cterm(a, kEyx, ox_, true, 0, nx, i, 0, v);
      }
      vy = LOSSY ? ca * vy + cb * (u - v) : vy + cb * (u - v);
    }
    if (xin && yin) {
      float u = hy - hyi, v = hx - hxj;
      if (MODE == kCpml) {
        if (xs) u = cterm(a, kEzx, ox_, true, 0, nx, i, 0, u);
        if (ys) v = cterm(a, kEzy, oy_, true, 1, ny, j, 0, v);
      }
      vz = LOSSY ? ca * vz + cb * (u - v) : vz + cb * (u - v);
    }
    if (MODE == kSponge) {
      const float sig = (sxy + a.sponge[nx + ny + k]) / 3.0f;
      if (sig != 0.f) {
        const float d = expf(((-sig) * a.sdt) / a.seps0);
        vx *= d;
        vy *= d;
        vz *= d;
      }
    }
    a.fo[0][c] = vx;
    a.fo[1][c] = vy;
    a.fo[2][c] = vz;
    hxm = hx;
    hym = hy;
  }
}

// ---------------------------------------------------------------------------
// Variant 2: the barrier-free recompute design of variant 1,
// reorganised so that every plane is straight-line code:
//  * CPML work is chosen per plane by a block-uniform mask -- bit 0: the
//    tile (with its x-1 column) touches an x slab, bit 1: a y slab, bit 2:
//    the plane lies in a z slab.  Each mask value is its own instantiation,
//    so interior planes carry no psi code at all, and slab planes no
//    per-cell branches: psi of a cell outside its slab is read at a clamped
//    offset, not selected and not written.
//  * Ex, Ey of the column, Ex of column x-1 and Ey of row y-1 at plane k+1
//    are loaded once and carried to the next plane in registers.
//  * 1-D launch in the XCD-aware tile order of Tiling.
// ---------------------------------------------------------------------------

// psi recursion + curl term of one cell, branch-free: when the cell is not
// in the slab (`in` false) psi is read at offset `base` (a valid slot
// element), the result is `d`, and nothing is written.
__device__ __forceinline__ float v2_psi(const float* __restrict__ pin, float* __restrict__ pout,
                                        unsigned base, unsigned off, bool in, bool wr,
                                        const float* __restrict__ tab, int n, int v, int kind,
                                        float d) {
  const float b = tab[(3 * kind + 0) * n + v];
  const float c = tab[(3 * kind + 1) * n + v];
  const float kinv = tab[(3 * kind + 2) * n + v];
  const unsigned o = base + (in ? off : 0u);
  const float ps = b * pin[o] + c * d;
  if (wr && in) put(&pout[o], ps);
  return in ? d * kinv + ps : d;
}

// per-column constants of variant 2
struct V2Col {
  int i, j, im, jm;            // column and its x-1 / y-1 neighbours (clamped)
  unsigned xp, xm, yp, ym;     // neighbour offsets, 0 on a non-periodic face
  bool xinH, yinH;             // H updates with an x+1 / y+1 neighbour
  bool xinE, yinE;             // E updates with an x-1 / y-1 neighbour
  // CPML: slab membership and slab index of i, i-1 (x) and j, j-1 (y)
  bool xs, xsm, ys, ysm;
  int sx, sxm, sy, sym;
  float sxy;                   // sponge: sigma_x[i] + sigma_y[j]
};

// values carried from plane k-1 to plane k
struct V2Carry {
  float ex, ey, exm, eym;  // Ex, Ey of the column, Ex(i-1), Ey(j-1) at plane k
  float hxm, hym;          // H^{n+1/2} x, y at plane k-1
};

template <int MODE, bool LOSSY, int SL>
__device__ __forceinline__ void v2_plane(
    const FusedArgs& a, const float* __restrict__ Ex, const float* __restrict__ Ey,
    const float* __restrict__ Ez, const float* __restrict__ Hx, const float* __restrict__ Hy,
    const float* __restrict__ Hz, float* __restrict__ oEx, float* __restrict__ oEy,
    float* __restrict__ oEz, float* __restrict__ oHx, float* __restrict__ oHy,
    float* __restrict__ oHz, const V2Col& q, int k, unsigned c, V2Carry& r) {
  constexpr bool per = MODE == kPeriodic;
  constexpr bool XS = MODE == kCpml && (SL & 1);
  constexpr bool YS = MODE == kCpml && (SL & 2);
  constexpr bool ZS = MODE == kCpml && (SL & 4);
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  const unsigned P = (unsigned)nx * ny;
  const bool zinH = per || k < nz - 1;
  const bool zinE = per || k > 0;
  const unsigned zp = (per && k == nz - 1) ? -(unsigned)(nz - 1) * P : (zinH ? P : 0u);
  const float ch = a.ch;

  // ---- loads: the column, its x+1 / y+1 / z+1 neighbours, and the inputs
  // of H at (i-1, j, k) and (i, j-1, k)
  const float ex = r.ex, ey = r.ey, exm = r.exm, eym = r.eym;
  const float ez = Ez[c];
  const float ey_xp = Ey[c + q.xp], ez_xp = Ez[c + q.xp];
  const float ex_yp = Ex[c + q.yp], ez_yp = Ez[c + q.yp];
  const float ex_zp = Ex[c + zp], ey_zp = Ey[c + zp];
  const float hx0 = Hx[c], hy0 = Hy[c], hz0 = Hz[c];
  const unsigned cxm = c + q.xm, cym = c + q.ym;
  const float ez_xm = Ez[cxm], ey_xm = Ey[cxm], ex_xm_yp = Ex[cxm + q.yp],
              ex_xm_zp = Ex[cxm + zp];
  const float hy_xm0 = Hy[cxm], hz_xm0 = Hz[cxm];
  const float ez_ym = Ez[cym], ex_ym = Ex[cym], ey_ym_xp = Ey[cym + q.xp],
              ey_ym_zp = Ey[cym + zp];
  const float hx_ym0 = Hx[cym], hz_ym0 = Hz[cym];
  const float cb = a.cb[c];
  const float ca = LOSSY ? a.ca[c] : 1.f;

  // CPML addressing (only used by the XS / YS / ZS instantiations)
  const float* tx = a.cpml;
  const float* ty = a.cpml + 6 * nx;
  const float* tz = a.cpml + 6 * (nx + ny);
  const unsigned ysl = (unsigned)a.ysl, zsl = (unsigned)a.zsl, xsl = (unsigned)a.xsl;
  const unsigned by0 = 0, bz0 = 4 * ysl, bx0 = 4 * ysl + 4 * zsl;
  const bool zs = ZS && (k < p || k >= nz - p);
  const int sz = zs ? slab_index(k, nz, p) : 0;
  const unsigned tp = 2u * (unsigned)p;
  // y group (k, s, i) ; z group (s, j, i) ; x group (k, j, s)
  auto offy = [&](int s, int ii) { return ((unsigned)k * tp + (unsigned)s) * nx + ii; };
  auto offz = [&](int jj, int ii) { return ((unsigned)sz * ny + jj) * nx + ii; };
  auto offx = [&](int jj, int s) { return ((unsigned)k * ny + jj) * tp + (unsigned)s; };
  const float* __restrict__ pin = a.pin;
  float* __restrict__ pout = a.pout;

  // ---- H^{n+1/2} of the column (fdtd3d.py:303-330)
  float hx, hy, hz;
  {
    const bool upd = q.yinH && zinH;
    float u = ez_yp - ez, v = ey_zp - ey;
    if (YS) u = v2_psi(pin, pout, by0 + kHxy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 1, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kHxz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 1, v);
    hx = upd ? hx0 - ch * (u - v) : hx0;
  }
  {
    const bool upd = q.xinH && zinH;
    float u = ex_zp - ex, v = ez_xp - ez;
    if (ZS) u = v2_psi(pin, pout, bz0 + (kHyz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 1, u);
    if (XS) v = v2_psi(pin, pout, bx0 + (kHyx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 1, v);
    hy = upd ? hy0 - ch * (u - v) : hy0;
  }
  {
    const bool upd = q.xinH && q.yinH;
    float u = ey_xp - ey, v = ex_yp - ex;
    if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 1, u);
    if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 1, v);
    hz = upd ? hz0 - ch * (u - v) : hz0;
  }
  // ---- Hy, Hz at (i-1, j, k) (recomputed, psi not written)
  float hy_xm, hz_xm;
  {
    float u = ex_xm_zp - exm, v = ez - ez_xm;
    if (ZS) u = v2_psi(pin, pout, bz0 + (kHyz - 4) * zsl, offz(q.j, q.im), zs, false, tz, nz, k, 1, u);
    if (XS) v = v2_psi(pin, pout, bx0 + (kHyx - 8) * xsl, offx(q.j, q.sxm), q.xsm, false, tx, nx, q.im, 1, v);
    hy_xm = zinH ? hy_xm0 - ch * (u - v) : hy_xm0;
  }
  {
    float u = ey - ey_xm, v = ex_xm_yp - exm;
    if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.j, q.sxm), q.xsm, false, tx, nx, q.im, 1, u);
    if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sy, q.im), q.ys, false, ty, ny, q.j, 1, v);
    hz_xm = q.yinH ? hz_xm0 - ch * (u - v) : hz_xm0;
  }
  // ---- Hx, Hz at (i, j-1, k)
  float hx_ym, hz_ym;
  {
    float u = ez - ez_ym, v = ey_ym_zp - eym;
    if (YS) u = v2_psi(pin, pout, by0 + kHxy * ysl, offy(q.sym, q.i), q.ysm, false, ty, ny, q.jm, 1, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kHxz - 4) * zsl, offz(q.jm, q.i), zs, false, tz, nz, k, 1, v);
    hx_ym = zinH ? hx_ym0 - ch * (u - v) : hx_ym0;
  }
  {
    float u = ey_ym_xp - eym, v = ex - ex_ym;
    if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.jm, q.sx), q.xs, false, tx, nx, q.i, 1, u);
    if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sym, q.i), q.ysm, false, ty, ny, q.jm, 1, v);
    hz_ym = q.xinH ? hz_ym0 - ch * (u - v) : hz_ym0;
  }
  put(&oHx[c], hx);
  put(&oHy[c], hy);
  put(&oHz[c], hz);

  // ---- E^{n+1} of the column (fdtd3d.py:332-359)
  float vx = ex, vy = ey, vz = ez;
  {
    const bool upd = q.yinE && zinE;
    float u = hz - hz_ym, v = hy - r.hym;
    if (YS) u = v2_psi(pin, pout, by0 + kExy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 0, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kExz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 0, v);
    const float w = LOSSY ? ca * ex + cb * (u - v) : ex + cb * (u - v);
    vx = upd ? w : ex;
  }
  {
    const bool upd = q.xinE && zinE;
    float u = hx - r.hxm, v = hz - hz_xm;
    if (ZS) u = v2_psi(pin, pout, bz0 + (kEyz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 0, u);
    if (XS) v = v2_psi(pin, pout, bx0 + (kEyx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 0, v);
    const float w = LOSSY ? ca * ey + cb * (u - v) : ey + cb * (u - v);
    vy = upd ? w : ey;
  }
  {
    const bool upd = q.xinE && q.yinE;
    float u = hy - hy_xm, v = hx - hx_ym;
    if (XS) u = v2_psi(pin, pout, bx0 + (kEzx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 0, u);
    if (YS) v = v2_psi(pin, pout, by0 + kEzy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 0, v);
    const float w = LOSSY ? ca * ez + cb * (u - v) : ez + cb * (u - v);
    vz = upd ? w : ez;
  }
  if (MODE == kSponge) {  // fdtd3d.py:361-366; a factor exp(-0) == 1 changes no bit
    const float sig = (q.sxy + a.sponge[nx + ny + k]) / 3.0f;
    const float d = sig != 0.f ? expf(((-sig) * a.sdt) / a.seps0) : 1.f;
    vx *= d;
    vy *= d;
    vz *= d;
  }
  put(&oEx[c], vx);
  put(&oEy[c], vy);
  put(&oEz[c], vz);

  r.ex = ex_zp;
  r.ey = ey_zp;
  r.exm = ex_xm_zp;
  r.eym = ey_ym_zp;
  r.hxm = hx;
  r.hym = hy;
}

// ---------------------------------------------------------------------------
// Variant 3 (default): variant 2 with the x neighbours taken from the
// neighbouring lanes.  A wavefront holds 64 consecutive x cells of one row,
// so E of column x+1 and the new H of column x-1 are already in registers
// of lane+1 / lane-1 and are fetched with a cross-lane permute (ds_bpermute,
// no LDS storage, no barrier).  Only the wave's edge lanes touch memory for
// them: lane 0 recomputes Hy, Hz of column x-1 (the previous tile's last
// column) from loads, and the last active lane loads E of column x+1.  Per
// plane a wave issues 15 full-width loads instead of 24.
// ---------------------------------------------------------------------------

// The loads of one plane of variant 3 (edge-lane loads exec-masked).
struct V3In {
  float ez, ex_yp, ez_yp, ex_zp, ey_zp, hx0, hy0, hz0;
  float ez_ym, ex_ym, ey_ym_zp, hx_ym0, hz_ym0, cb, ca;
  unsigned mid;  // material id (palette)
  float ld_ey_xp, ld_ez_xp, ld_ey_ym_xp;
  float ez_xm, ey_xm, ex_xm_yp, ex_xm_zp, hy_xm0, hz_xm0;
};

template <int MODE, bool LOSSY>
__device__ __forceinline__ unsigned v3_zp(const FusedArgs& a, int k) {
  const unsigned P = (unsigned)a.nx * a.ny;
  if (MODE == kPeriodic) return k == a.nz - 1 ? -(unsigned)(a.nz - 1) * P : P;
  return k < a.nz - 1 ? P : 0u;
}

template <int MODE, bool LOSSY>
__device__ __forceinline__ void v3_load(
    const FusedArgs& a, const float* __restrict__ Ex, const float* __restrict__ Ey,
    const float* __restrict__ Ez, const float* __restrict__ Hx, const float* __restrict__ Hy,
    const float* __restrict__ Hz, const V2Col& q, bool lane0, bool edge_r, int k, unsigned c,
    V3In& L) {
  const unsigned zp = v3_zp<MODE, LOSSY>(a, k);
  // full-width loads: the column, its y+1 / z+1 neighbours, row y-1
  L.ez = Ez[c];
  L.ex_yp = Ex[c + q.yp];
  L.ez_yp = Ez[c + q.yp];
  L.ex_zp = Ex[c + zp];
  L.ey_zp = Ey[c + zp];
  L.hx0 = Hx[c];
  L.hy0 = Hy[c];
  L.hz0 = Hz[c];
  const unsigned cym = c + q.ym;
  L.ez_ym = Ez[cym];
  L.ex_ym = Ex[cym];
  L.ey_ym_zp = Ey[cym + zp];
  L.hx_ym0 = Hx[cym];
  L.hz_ym0 = Hz[cym];
  if (a.mat) {
    L.mid = a.mat[c];
  } else {
    L.cb = a.cb[c];
    L.ca = LOSSY ? a.ca[c] : 1.f;
  }
  // edge lanes (exec-masked)
  L.ld_ey_xp = L.ld_ez_xp = L.ld_ey_ym_xp = 0.f;
  if (edge_r) {
    L.ld_ey_xp = Ey[c + q.xp];
    L.ld_ez_xp = Ez[c + q.xp];
    L.ld_ey_ym_xp = Ey[cym + q.xp];
  }
  const unsigned cxm = c + q.xm;
  L.ez_xm = L.ey_xm = L.ex_xm_yp = L.ex_xm_zp = L.hy_xm0 = L.hz_xm0 = 0.f;
  if (lane0) {
    L.ez_xm = Ez[cxm];
    L.ey_xm = Ey[cxm];
    L.ex_xm_yp = Ex[cxm + q.yp];
    L.ex_xm_zp = Ex[cxm + zp];
    L.hy_xm0 = Hy[cxm];
    L.hz_xm0 = Hz[cxm];
  }
}

// Arithmetic and stores of one plane of variant 3 from its loads.
template <int MODE, bool LOSSY, int SL>
__device__ __forceinline__ void v3_compute(
    const FusedArgs& a, float* __restrict__ oEx, float* __restrict__ oEy,
    float* __restrict__ oEz, float* __restrict__ oHx, float* __restrict__ oHy,
    float* __restrict__ oHz, const V2Col& q, bool lane0, bool edge_r, int k, unsigned c,
    const V3In& L, V2Carry& r, const float* spal) {
  constexpr bool per = MODE == kPeriodic;
  constexpr bool XS = MODE == kCpml && (SL & 1);
  constexpr bool YS = MODE == kCpml && (SL & 2);
  constexpr bool ZS = MODE == kCpml && (SL & 4);
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  const bool zinH = per || k < nz - 1;
  const bool zinE = per || k > 0;
  const float ch = a.ch;
  const float ex = r.ex, ey = r.ey, eym = r.eym;
  const float ez = L.ez, ex_yp = L.ex_yp, ez_yp = L.ez_yp, ex_zp = L.ex_zp, ey_zp = L.ey_zp;
  const float hx0 = L.hx0, hy0 = L.hy0, hz0 = L.hz0;
  const float ez_ym = L.ez_ym, ex_ym = L.ex_ym, ey_ym_zp = L.ey_ym_zp;
  const float hx_ym0 = L.hx_ym0, hz_ym0 = L.hz_ym0;
  const float cb = a.mat ? spal[L.mid] : L.cb;
  const float ca = a.mat ? (LOSSY ? spal[256 + L.mid] : 1.f) : L.ca;
  const float ld_ey_xp = L.ld_ey_xp, ld_ez_xp = L.ld_ez_xp, ld_ey_ym_xp = L.ld_ey_ym_xp;
  const float ez_xm = L.ez_xm, ey_xm = L.ey_xm, ex_xm_yp = L.ex_xm_yp, ex_xm_zp = L.ex_xm_zp;
  const float hy_xm0 = L.hy_xm0, hz_xm0 = L.hz_xm0;
  // ---- column x+1 from lane+1 (the last active lane uses its loads)
  float ey_xp = __shfl_down(ey, 1);
  float ez_xp = __shfl_down(ez, 1);
  float ey_ym_xp = __shfl_down(eym, 1);
  if (edge_r) {
    ey_xp = ld_ey_xp;
    ez_xp = ld_ez_xp;
    ey_ym_xp = ld_ey_ym_xp;
  }

  const float* tx = a.cpml;
  const float* ty = a.cpml + 6 * nx;
  const float* tz = a.cpml + 6 * (nx + ny);
  const unsigned ysl = (unsigned)a.ysl, zsl = (unsigned)a.zsl, xsl = (unsigned)a.xsl;
  const unsigned by0 = 0, bz0 = 4 * ysl, bx0 = 4 * ysl + 4 * zsl;
  const bool zs = ZS && (k < p || k >= nz - p);
  const int sz = zs ? slab_index(k, nz, p) : 0;
  const unsigned tp = 2u * (unsigned)p;
  auto offy = [&](int s, int ii) { return ((unsigned)k * tp + (unsigned)s) * nx + ii; };
  auto offz = [&](int jj, int ii) { return ((unsigned)sz * ny + jj) * nx + ii; };
  auto offx = [&](int jj, int s) { return ((unsigned)k * ny + jj) * tp + (unsigned)s; };
  const float* __restrict__ pin = a.pin;
  float* __restrict__ pout = a.pout;

  // ---- H^{n+1/2} of the column (fdtd3d.py:303-330)
  float hx, hy, hz;
  {
    const bool upd = q.yinH && zinH;
    float u = ez_yp - ez, v = ey_zp - ey;
    if (YS) u = v2_psi(pin, pout, by0 + kHxy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 1, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kHxz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 1, v);
    hx = upd ? hx0 - ch * (u - v) : hx0;
  }
  {
    const bool upd = q.xinH && zinH;
    float u = ex_zp - ex, v = ez_xp - ez;
    if (ZS) u = v2_psi(pin, pout, bz0 + (kHyz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 1, u);
    if (XS) v = v2_psi(pin, pout, bx0 + (kHyx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 1, v);
    hy = upd ? hy0 - ch * (u - v) : hy0;
  }
  {
    const bool upd = q.xinH && q.yinH;
    float u = ey_xp - ey, v = ex_yp - ex;
    if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 1, u);
    if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 1, v);
    hz = upd ? hz0 - ch * (u - v) : hz0;
  }
  // ---- Hy, Hz of column x-1: lane-1's new values; lane 0 recomputes them
  float hy_xm = __shfl_up(hy, 1);
  float hz_xm = __shfl_up(hz, 1);
  if (lane0) {
    const float exm = r.exm;
    {
      float u = ex_xm_zp - exm, v = ez - ez_xm;
      if (ZS) u = v2_psi(pin, pout, bz0 + (kHyz - 4) * zsl, offz(q.j, q.im), zs, false, tz, nz, k, 1, u);
      if (XS) v = v2_psi(pin, pout, bx0 + (kHyx - 8) * xsl, offx(q.j, q.sxm), q.xsm, false, tx, nx, q.im, 1, v);
      hy_xm = zinH ? hy_xm0 - ch * (u - v) : hy_xm0;
    }
    {
      float u = ey - ey_xm, v = ex_xm_yp - exm;
      if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.j, q.sxm), q.xsm, false, tx, nx, q.im, 1, u);
      if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sy, q.im), q.ys, false, ty, ny, q.j, 1, v);
      hz_xm = q.yinH ? hz_xm0 - ch * (u - v) : hz_xm0;
    }
  }
  r.exm = ex_xm_zp;  // meaningful in lane 0 only
  // ---- Hx, Hz at (i, j-1, k)
  float hx_ym, hz_ym;
  {
    float u = ez - ez_ym, v = ey_ym_zp - eym;
    if (YS) u = v2_psi(pin, pout, by0 + kHxy * ysl, offy(q.sym, q.i), q.ysm, false, ty, ny, q.jm, 1, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kHxz - 4) * zsl, offz(q.jm, q.i), zs, false, tz, nz, k, 1, v);
    hx_ym = zinH ? hx_ym0 - ch * (u - v) : hx_ym0;
  }
  {
    float u = ey_ym_xp - eym, v = ex - ex_ym;
    if (XS) u = v2_psi(pin, pout, bx0 + (kHzx - 8) * xsl, offx(q.jm, q.sx), q.xs, false, tx, nx, q.i, 1, u);
    if (YS) v = v2_psi(pin, pout, by0 + kHzy * ysl, offy(q.sym, q.i), q.ysm, false, ty, ny, q.jm, 1, v);
    hz_ym = q.xinH ? hz_ym0 - ch * (u - v) : hz_ym0;
  }
  put(&oHx[c], hx);
  put(&oHy[c], hy);
  put(&oHz[c], hz);

  // ---- E^{n+1} of the column (fdtd3d.py:332-359)
  float vx, vy, vz;
  {
    const bool upd = q.yinE && zinE;
    float u = hz - hz_ym, v = hy - r.hym;
    if (YS) u = v2_psi(pin, pout, by0 + kExy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 0, u);
    if (ZS) v = v2_psi(pin, pout, bz0 + (kExz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 0, v);
    const float w = LOSSY ? ca * ex + cb * (u - v) : ex + cb * (u - v);
    vx = upd ? w : ex;
  }
  {
    const bool upd = q.xinE && zinE;
    float u = hx - r.hxm, v = hz - hz_xm;
    if (ZS) u = v2_psi(pin, pout, bz0 + (kEyz - 4) * zsl, offz(q.j, q.i), zs, upd, tz, nz, k, 0, u);
    if (XS) v = v2_psi(pin, pout, bx0 + (kEyx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 0, v);
    const float w = LOSSY ? ca * ey + cb * (u - v) : ey + cb * (u - v);
    vy = upd ? w : ey;
  }
  {
    const bool upd = q.xinE && q.yinE;
    float u = hy - hy_xm, v = hx - hx_ym;
    if (XS) u = v2_psi(pin, pout, bx0 + (kEzx - 8) * xsl, offx(q.j, q.sx), q.xs, upd, tx, nx, q.i, 0, u);
    if (YS) v = v2_psi(pin, pout, by0 + kEzy * ysl, offy(q.sy, q.i), q.ys, upd, ty, ny, q.j, 0, v);
    const float w = LOSSY ? ca * ez + cb * (u - v) : ez + cb * (u - v);
    vz = upd ? w : ez;
  }
  if (MODE == kSponge) {  // fdtd3d.py:361-366; a factor exp(-0) == 1 changes no bit
    const float sig = (q.sxy + a.sponge[nx + ny + k]) / 3.0f;
    const float d = sig != 0.f ? expf(((-sig) * a.sdt) / a.seps0) : 1.f;
    vx *= d;
    vy *= d;
    vz *= d;
  }
  put(&oEx[c], vx);
  put(&oEy[c], vy);
  put(&oEz[c], vz);

  r.ex = ex_zp;
  r.ey = ey_zp;
  r.eym = ey_ym_zp;
  r.hxm = hx;
  r.hym = hy;
}

template <int MODE, bool LOSSY, int SL>
__device__ __forceinline__ void v3_plane(
    const FusedArgs& a, const float* __restrict__ Ex, const float* __restrict__ Ey,
    const float* __restrict__ Ez, const float* __restrict__ Hx, const float* __restrict__ Hy,
    const float* __restrict__ Hz, float* __restrict__ oEx, float* __restrict__ oEy,
    float* __restrict__ oEz, float* __restrict__ oHx, float* __restrict__ oHy,
    float* __restrict__ oHz, const V2Col& q, bool lane0, bool edge_r, int k, unsigned c,
    V2Carry& r, const float* spal) {
  V3In L;
  v3_load<MODE, LOSSY>(a, Ex, Ey, Ez, Hx, Hy, Hz, q, lane0, edge_r, k, c, L);
  v3_compute<MODE, LOSSY, SL>(a, oEx, oEy, oEz, oHx, oHy, oHz, q, lane0, edge_r, k, c, L, r,
                              spal);
}

// ---------------------------------------------------------------------------
// Variant 3, CPML slab shell.  Same arithmetic as v3_plane with every psi
// term, but the y and z psi work sits behind wave-uniform branches (j is
// one value per wavefront, k one per plane), the x table values are loaded
// once per tile, and only the
// x terms are selected per lane (their psi loads are issued with the field
// loads).
// ---------------------------------------------------------------------------

// CPML x coefficients of a column (per lane, loaded once per tile)
struct V3Tab {
  float xbE, xcE, xkE, xbH, xcH, xkH;
};

// psi' = b psi + c d (returned in ps) and the curl term d / kappa + psi'
__device__ __forceinline__ float pterm(float b, float c, float kinv, float psi, float d,
                                       float& ps) {
  ps = b * psi + c * d;
  return d * kinv + ps;
}

template <int MODE, bool LOSSY, bool XS>
__device__ __forceinline__ void v3s_plane(
    const FusedArgs& a, const float* __restrict__ Ex, const float* __restrict__ Ey,
    const float* __restrict__ Ez, const float* __restrict__ Hx, const float* __restrict__ Hy,
    const float* __restrict__ Hz, float* __restrict__ oEx, float* __restrict__ oEy,
    float* __restrict__ oEz, float* __restrict__ oHx, float* __restrict__ oHy,
    float* __restrict__ oHz, const V2Col& q, const V3Tab& t, bool lane0, bool edge_r, int k,
    unsigned c, V2Carry& r, const float* spal) {
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  const unsigned P = (unsigned)nx * ny;
  const bool zinH = k < nz - 1;
  const bool zinE = k > 0;
  const unsigned zp = zinH ? P : 0u;
  const float ch = a.ch;

  // ---- field loads (as v3_plane)
  const float ex = r.ex, ey = r.ey, eym = r.eym;
  const float ez = Ez[c];
  const float ex_yp = Ex[c + q.yp], ez_yp = Ez[c + q.yp];
  const float ex_zp = Ex[c + zp], ey_zp = Ey[c + zp];
  const float hx0 = Hx[c], hy0 = Hy[c], hz0 = Hz[c];
  const unsigned cym = c + q.ym;
  const float ez_ym = Ez[cym], ex_ym = Ex[cym], ey_ym_zp = Ey[cym + zp];
  const float hx_ym0 = Hx[cym], hz_ym0 = Hz[cym];
  // coefficients: per-cell arrays, or a material id looked up in the LDS
  // table at the E update
  float cb = 0.f, ca = 1.f;
  unsigned mid = 0;
  if (a.mat) {
    mid = a.mat[c];
  } else {
    cb = a.cb[c];
    ca = LOSSY ? a.ca[c] : 1.f;
  }
  float ld_ey_xp = 0.f, ld_ez_xp = 0.f, ld_ey_ym_xp = 0.f;
  if (edge_r) {
    ld_ey_xp = Ey[c + q.xp];
    ld_ez_xp = Ez[c + q.xp];
    ld_ey_ym_xp = Ey[cym + q.xp];
  }
  const unsigned cxm = c + q.xm;
  float ez_xm = 0.f, ey_xm = 0.f, ex_xm_yp = 0.f, ex_xm_zp = 0.f, hy_xm0 = 0.f, hz_xm0 = 0.f;
  if (lane0) {
    ez_xm = Ez[cxm];
    ey_xm = Ey[cxm];
    ex_xm_yp = Ex[cxm + q.yp];
    ex_xm_zp = Ex[cxm + zp];
    hy_xm0 = Hy[cxm];
    hz_xm0 = Hz[cxm];
  }

  // ---- psi addressing: slot bases and offsets within a slot
  const float* __restrict__ pin = a.pin;
  float* __restrict__ pout = a.pout;
  const unsigned ysl = (unsigned)a.ysl, zsl = (unsigned)a.zsl, xsl = (unsigned)a.xsl;
  const unsigned bz0 = 4 * ysl, bx0 = 4 * ysl + 4 * zsl;
  const unsigned tp = 2u * (unsigned)p;
  const bool zs = k < p || k >= nz - p;
  const int sz = zs ? slab_index(k, nz, p) : 0;
  const unsigned oy = ((unsigned)k * tp + (unsigned)q.sy) * nx + q.i;
  const unsigned oym = ((unsigned)k * tp + (unsigned)q.sym) * nx + q.i;
  const unsigned oz = ((unsigned)sz * ny + q.j) * nx + q.i;
  const unsigned ozm = ((unsigned)sz * ny + q.jm) * nx + q.i;
  const unsigned ox = q.xs ? ((unsigned)k * ny + q.j) * tp + (unsigned)q.sx : 0u;
  const unsigned oxm = q.xs ? ((unsigned)k * ny + q.jm) * tp + (unsigned)q.sx : 0u;

  // ---- x psi loads (per lane, clamped); y and z psi are loaded inside
  // their wave-uniform branches
  const float* ty = a.cpml + 6 * nx;
  const float* tz = a.cpml + 6 * (nx + ny);
  float pHyx = 0.f, pHzx = 0.f, pEyx = 0.f, pEzx = 0.f, pHzx_m = 0.f;
  if (XS) {
    pHyx = pin[bx0 + (kHyx - 8) * xsl + ox];
    pHzx = pin[bx0 + (kHzx - 8) * xsl + ox];
    pEyx = pin[bx0 + (kEyx - 8) * xsl + ox];
    pEzx = pin[bx0 + (kEzx - 8) * xsl + ox];
    pHzx_m = pin[bx0 + (kHzx - 8) * xsl + oxm];
  }
  // y / z psi of the column and of row y-1, loaded up front in the
  // instantiation without x terms (its register budget allows it)
  float eHxy = 0.f, eHzy = 0.f, eExy = 0.f, eEzy = 0.f, eHxy_m = 0.f, eHzy_m = 0.f;
  float eHxz = 0.f, eHyz = 0.f, eExz = 0.f, eEyz = 0.f, eHxz_m = 0.f;
  if constexpr (!XS) {
    if (q.ys) {
      eHxy = pin[kHxy * ysl + oy];
      eHzy = pin[kHzy * ysl + oy];
      eExy = pin[kExy * ysl + oy];
      eEzy = pin[kEzy * ysl + oy];
    }
    if (q.ysm) {
      eHxy_m = pin[kHxy * ysl + oym];
      eHzy_m = pin[kHzy * ysl + oym];
    }
    if (zs) {
      eHxz = pin[bz0 + (kHxz - 4) * zsl + oz];
      eHyz = pin[bz0 + (kHyz - 4) * zsl + oz];
      eExz = pin[bz0 + (kExz - 4) * zsl + oz];
      eEyz = pin[bz0 + (kEyz - 4) * zsl + oz];
      eHxz_m = pin[bz0 + (kHxz - 4) * zsl + ozm];
    }
  }
  // psi value: the early load (no x terms) or a load at the point of use
#define YEE_PSI(early, off) (XS ? pin[off] : (early))
  // lane 0: psi and x tables of column x-1
  float l_pHyx = 0.f, l_pHzx = 0.f;
  float lxbH = 0.f, lxcH = 0.f, lxkH = 0.f;
  if (lane0) {
    if (XS && q.xsm) {
      const unsigned oxl = ((unsigned)k * ny + q.j) * tp + (unsigned)q.sxm;
      l_pHyx = pin[bx0 + (kHyx - 8) * xsl + oxl];
      l_pHzx = pin[bx0 + (kHzx - 8) * xsl + oxl];
      lxbH = a.cpml[3 * nx + q.im];
      lxcH = a.cpml[4 * nx + q.im];
      lxkH = a.cpml[5 * nx + q.im];
    }
  }
  float ey_xp = __shfl_down(ey, 1);
  float ez_xp = __shfl_down(ez, 1);
  float ey_ym_xp = __shfl_down(eym, 1);
  if (edge_r) {
    ey_xp = ld_ey_xp;
    ez_xp = ld_ez_xp;
    ey_ym_xp = ld_ey_ym_xp;
  }

  // ---- H^{n+1/2} of the column (fdtd3d.py:303-330)
  float hx, hy, hz, ps;
  {
    const bool upd = q.yinH && zinH;
    float u = ez_yp - ez, v = ey_zp - ey;
    if (q.ys) {
      u = pterm(ty[3 * ny + q.j], ty[4 * ny + q.j], ty[5 * ny + q.j], YEE_PSI(eHxy, kHxy * ysl + oy), u, ps);
      if (upd) put(&pout[kHxy * ysl + oy], ps);
    }
    if (zs) {
      v = pterm(tz[3 * nz + k], tz[4 * nz + k], tz[5 * nz + k], YEE_PSI(eHxz, bz0 + (kHxz - 4) * zsl + oz), v, ps);
      if (upd) put(&pout[bz0 + (kHxz - 4) * zsl + oz], ps);
    }
    hx = upd ? hx0 - ch * (u - v) : hx0;
  }
  {
    const bool upd = q.xinH && zinH;
    float u = ex_zp - ex, v = ez_xp - ez;
    if (zs) {
      u = pterm(tz[3 * nz + k], tz[4 * nz + k], tz[5 * nz + k], YEE_PSI(eHyz, bz0 + (kHyz - 4) * zsl + oz), u, ps);
      if (upd) put(&pout[bz0 + (kHyz - 4) * zsl + oz], ps);
    }
    if (XS) {
      const float w = pterm(t.xbH, t.xcH, t.xkH, pHyx, v, ps);
      if (upd && q.xs) put(&pout[bx0 + (kHyx - 8) * xsl + ox], ps);
      v = q.xs ? w : v;
    }
    hy = upd ? hy0 - ch * (u - v) : hy0;
  }
  {
    const bool upd = q.xinH && q.yinH;
    float u = ey_xp - ey, v = ex_yp - ex;
    if (XS) {
      const float w = pterm(t.xbH, t.xcH, t.xkH, pHzx, u, ps);
      if (upd && q.xs) put(&pout[bx0 + (kHzx - 8) * xsl + ox], ps);
      u = q.xs ? w : u;
    }
    if (q.ys) {
      v = pterm(ty[3 * ny + q.j], ty[4 * ny + q.j], ty[5 * ny + q.j], YEE_PSI(eHzy, kHzy * ysl + oy), v, ps);
      if (upd) put(&pout[kHzy * ysl + oy], ps);
    }
    hz = upd ? hz0 - ch * (u - v) : hz0;
  }
  // ---- Hy, Hz of column x-1: lane-1's new values; lane 0 recomputes them
  float hy_xm = __shfl_up(hy, 1);
  float hz_xm = __shfl_up(hz, 1);
  if (lane0) {
    const float exm = r.exm;
    {
      float u = ex_xm_zp - exm, v = ez - ez_xm;
      if (zs) u = pterm(tz[3 * nz + k], tz[4 * nz + k], tz[5 * nz + k], pin[bz0 + (kHyz - 4) * zsl + oz + q.xm], u, ps);
      if (XS && q.xsm) v = pterm(lxbH, lxcH, lxkH, l_pHyx, v, ps);
      hy_xm = zinH ? hy_xm0 - ch * (u - v) : hy_xm0;
    }
    {
      float u = ey - ey_xm, v = ex_xm_yp - exm;
      if (XS && q.xsm) u = pterm(lxbH, lxcH, lxkH, l_pHzx, u, ps);
      if (q.ys) v = pterm(ty[3 * ny + q.j], ty[4 * ny + q.j], ty[5 * ny + q.j], pin[kHzy * ysl + oy + q.xm], v, ps);
      hz_xm = q.yinH ? hz_xm0 - ch * (u - v) : hz_xm0;
    }
  }
  r.exm = ex_xm_zp;  // meaningful in lane 0 only
  // ---- Hx, Hz at (i, j-1, k)
  float hx_ym, hz_ym;
  {
    float u = ez - ez_ym, v = ey_ym_zp - eym;
    if (q.ysm) u = pterm(ty[3 * ny + q.jm], ty[4 * ny + q.jm], ty[5 * ny + q.jm], YEE_PSI(eHxy_m, kHxy * ysl + oym), u, ps);
    if (zs) v = pterm(tz[3 * nz + k], tz[4 * nz + k], tz[5 * nz + k], YEE_PSI(eHxz_m, bz0 + (kHxz - 4) * zsl + ozm), v, ps);
    hx_ym = zinH ? hx_ym0 - ch * (u - v) : hx_ym0;
  }
  {
    float u = ey_ym_xp - eym, v = ex - ex_ym;
    if (XS) {
      const float w = pterm(t.xbH, t.xcH, t.xkH, pHzx_m, u, ps);
      u = q.xs ? w : u;
    }
    if (q.ysm) v = pterm(ty[3 * ny + q.jm], ty[4 * ny + q.jm], ty[5 * ny + q.jm], YEE_PSI(eHzy_m, kHzy * ysl + oym), v, ps);
    hz_ym = q.xinH ? hz_ym0 - ch * (u - v) : hz_ym0;
  }
  put(&oHx[c], hx);
  put(&oHy[c], hy);
  put(&oHz[c], hz);

  if (a.mat) {
    cb = spal[mid];
    ca = LOSSY ? spal[256 + mid] : 1.f;
  }
  // ---- E^{n+1} of the column (fdtd3d.py:332-359)
  float vx, vy, vz;
  {
    const bool upd = q.yinE && zinE;
    float u = hz - hz_ym, v = hy - r.hym;
    if (q.ys) {
      u = pterm(ty[q.j], ty[ny + q.j], ty[2 * ny + q.j], YEE_PSI(eExy, kExy * ysl + oy), u, ps);
      if (upd) put(&pout[kExy * ysl + oy], ps);
    }
    if (zs) {
      v = pterm(tz[k], tz[nz + k], tz[2 * nz + k], YEE_PSI(eExz, bz0 + (kExz - 4) * zsl + oz), v, ps);
      if (upd) put(&pout[bz0 + (kExz - 4) * zsl + oz], ps);
    }
    const float w = LOSSY ? ca * ex + cb * (u - v) : ex + cb * (u - v);
    vx = upd ? w : ex;
  }
  {
    const bool upd = q.xinE && zinE;
    float u = hx - r.hxm, v = hz - hz_xm;
    if (zs) {
      u = pterm(tz[k], tz[nz + k], tz[2 * nz + k], YEE_PSI(eEyz, bz0 + (kEyz - 4) * zsl + oz), u, ps);
      if (upd) put(&pout[bz0 + (kEyz - 4) * zsl + oz], ps);
    }
    if (XS) {
      const float w = pterm(t.xbE, t.xcE, t.xkE, pEyx, v, ps);
      if (upd && q.xs) put(&pout[bx0 + (kEyx - 8) * xsl + ox], ps);
      v = q.xs ? w : v;
    }
    const float w = LOSSY ? ca * ey + cb * (u - v) : ey + cb * (u - v);
    vy = upd ? w : ey;
  }
  {
    const bool upd = q.xinE && q.yinE;
    float u = hy - hy_xm, v = hx - hx_ym;
    if (XS) {
      const float w = pterm(t.xbE, t.xcE, t.xkE, pEzx, u, ps);
      if (upd && q.xs) put(&pout[bx0 + (kEzx - 8) * xsl + ox], ps);
      u = q.xs ? w : u;
    }
    if (q.ys) {
      v = pterm(ty[q.j], ty[ny + q.j], ty[2 * ny + q.j], YEE_PSI(eEzy, kEzy * ysl + oy), v, ps);
      if (upd) put(&pout[kEzy * ysl + oy], ps);
    }
    const float w = LOSSY ? ca * ez + cb * (u - v) : ez + cb * (u - v);
    vz = upd ? w : ez;
  }
  put(&oEx[c], vx);
  put(&oEy[c], vy);
  put(&oEz[c], vz);

  r.ex = ex_zp;
  r.ey = ey_zp;
  r.eym = ey_ym_zp;
  r.hxm = hx;
  r.hym = hy;
}
#undef YEE_PSI

// SHELL == 0 with CPML: the tiles of `ts` lie outside every slab (no psi
// code is compiled in).  SHELL == 1, 2: any tiles, slab work chosen per
// plane; with VAR >= 3, 1 = tiles whose columns (and x-1 neighbours) touch
// no x slab, 2 = x-slab columns (x psi terms compiled in).
template <int MODE, bool LOSSY, int SHELL, int VAR, int ROWS>
__global__ __launch_bounds__(64 * ROWS) void yee_fused_v2_kernel(
    FusedArgs a, TileSet ts, const float* __restrict__ fi0,
    const float* __restrict__ fi1, const float* __restrict__ fi2, const float* __restrict__ fi3,
    const float* __restrict__ fi4, const float* __restrict__ fi5, float* __restrict__ fo0,
    float* __restrict__ fo1, float* __restrict__ fo2, float* __restrict__ fo3,
    float* __restrict__ fo4, float* __restrict__ fo5) {
  TilePos tp;
  if (!tile_of_set(ts, tp)) return;  // (the whole workgroup)
  // material palette in LDS (variants 3-4)
  __shared__ float spal[512];
  if (VAR >= 3 && a.mat) {
    for (int q = threadIdx.x; q < 512; q += 64 * ROWS) spal[q] = a.pal[q];
    __syncthreads();
  }
  const int nx = a.nx, ny = a.ny, nz = a.nz, p = a.p;
  constexpr bool per = MODE == kPeriodic;
  const int k0 = tp.k0, k1 = tp.k1;
  const int lane = threadIdx.x & 63;
  // a paired tile holds two 32-column segments, one per half-wave; the
  // first / last active lane of each segment exchanges its x neighbour
  // through memory instead of a lane permute
  const int seg = tp.pair ? (lane & 31) : lane;
  const int i = tp.pair && lane >= 32 ? tp.ih0 + seg : tp.i0 + seg;
  const int xl = tp.pair && lane >= 32 ? tp.ihl : tp.xl;
  const bool lane0 = seg == 0;
  const bool edge_r = seg == (tp.pair ? 31 : 63) || i == xl - 1;
  // one y row per wavefront: j and everything derived from it is scalar
  const int j = __builtin_amdgcn_readfirstlane(tp.j0 + (int)(threadIdx.x >> 6));
  if (i >= xl || j >= tp.yl) return;
  V2Col q;
  q.i = i;
  q.j = j;
  q.xinH = per || i < nx - 1;
  q.yinH = per || j < ny - 1;
  q.xinE = per || i > 0;
  q.yinE = per || j > 0;
  q.im = i > 0 ? i - 1 : (per ? nx - 1 : i);
  q.jm = j > 0 ? j - 1 : (per ? ny - 1 : j);
  q.xp = i < nx - 1 ? 1u : (per ? -(unsigned)(nx - 1) : 0u);
  q.yp = j < ny - 1 ? (unsigned)nx : (per ? -(unsigned)(ny - 1) * nx : 0u);
  q.xm = (unsigned)q.im - (unsigned)i;
  q.ym = ((unsigned)q.jm - (unsigned)j) * (unsigned)nx;
  q.xs = MODE == kCpml && (i < p || i >= nx - p);
  q.xsm = MODE == kCpml && (q.im < p || q.im >= nx - p);
  q.ys = MODE == kCpml && (j < p || j >= ny - p);
  q.ysm = MODE == kCpml && (q.jm < p || q.jm >= ny - p);
  q.sx = q.xs ? slab_index(i, nx, p) : 0;
  q.sxm = q.xsm ? slab_index(q.im, nx, p) : 0;
  q.sy = q.ys ? slab_index(j, ny, p) : 0;
  q.sym = q.ysm ? slab_index(q.jm, ny, p) : 0;
  q.sxy = MODE == kSponge ? a.sponge[i] + a.sponge[nx + j] : 0.f;
  // block-uniform slab mask of the tile (its columns and their x-1 / y-1
  // neighbours)
  int slxy = 0;
  if (MODE == kCpml) {
    const int il = min(tp.i0 + 63, tp.xl - 1);
    const int jl = min(tp.j0 + ROWS - 1, tp.yl - 1);
    if (tp.pair || tp.i0 <= p || il >= nx - p) slxy |= 1;
    if (tp.j0 <= p || jl >= ny - p) slxy |= 2;
  }
  const unsigned P = (unsigned)nx * ny;
  unsigned c = (unsigned)k0 * P + (unsigned)j * nx + i;
  V2Carry r;
  r.ex = fi0[c];
  r.ey = fi1[c];
  r.exm = fi0[c + q.xm];
  r.eym = fi1[c + q.ym];
  r.hxm = 0.f;
  r.hym = 0.f;
  {
    int kp = k0 - 1;
    if (per && kp < 0) kp += nz;
    if (kp >= 0) {
      float unused;
      h_cell<MODE, 3>(a, i, j, kp, false, r.hxm, r.hym, unused, nullptr);
    }
  }
  V3Tab tb;
  if constexpr (MODE == kCpml && SHELL == 2 && VAR >= 3) {
    const float* tx = a.cpml;
    tb.xbE = tx[i]; tb.xcE = tx[nx + i]; tb.xkE = tx[2 * nx + i];
    tb.xbH = tx[3 * nx + i]; tb.xcH = tx[4 * nx + i]; tb.xkH = tx[5 * nx + i];
  }
#define YEE_V3S_PLANE(XSB)                                                                     \
  v3s_plane<MODE, LOSSY, XSB>(a, fi0, fi1, fi2, fi3, fi4, fi5, fo0, fo1, fo2, fo3, fo4, fo5, q, \
                              tb, lane0, edge_r, k, c, r, spal)
#define YEE_V2_PLANE(S)                                                                       \
  if constexpr (VAR >= 3)                                                                     \
    v3_plane<MODE, LOSSY, S>(a, fi0, fi1, fi2, fi3, fi4, fi5, fo0, fo1, fo2, fo3, fo4, fo5, q, \
                             lane0, edge_r, k, c, r, spal);                                   \
  else                                                                                        \
    v2_plane<MODE, LOSSY, S>(a, fi0, fi1, fi2, fi3, fi4, fi5, fo0, fo1, fo2, fo3, fo4, fo5, q, \
                             k, c, r)
  if constexpr (VAR == 4 && !(MODE == kCpml && SHELL != 0)) {
    // software pipeline: the loads of plane k+1 are issued before the
    // arithmetic and stores of plane k
    V3In cur, nxt;
    v3_load<MODE, LOSSY>(a, fi0, fi1, fi2, fi3, fi4, fi5, q, lane0, edge_r, k0, c, cur);
    for (int k = k0; k < k1; ++k, c += P) {
      if (k + 1 < k1)
        v3_load<MODE, LOSSY>(a, fi0, fi1, fi2, fi3, fi4, fi5, q, lane0, edge_r, k + 1, c + P, nxt);
      v3_compute<MODE, LOSSY, 0>(a, fo0, fo1, fo2, fo3, fo4, fo5, q, lane0, edge_r, k, c, cur, r,
                                 spal);
      cur = nxt;
    }
  } else
  for (int k = k0; k < k1; ++k, c += P) {
    if constexpr (MODE == kCpml && SHELL != 0 && VAR >= 3) {
      YEE_V3S_PLANE(SHELL == 2);
    } else if (MODE != kCpml || SHELL == 0) {
      YEE_V2_PLANE(0);
    } else {
      const int sl = slxy | ((k < p || k >= nz - p) ? 4 : 0);
      switch (sl) {
        case 0: YEE_V2_PLANE(0); break;
        case 1: YEE_V2_PLANE(1); break;
        case 2: YEE_V2_PLANE(2); break;
        case 3: YEE_V2_PLANE(3); break;
        case 4: YEE_V2_PLANE(4); break;
        case 5: YEE_V2_PLANE(5); break;
        case 6: YEE_V2_PLANE(6); break;
        default: YEE_V2_PLANE(7); break;
      }
    }
  }
#undef YEE_V2_PLANE
#undef YEE_V3S_PLANE
}

constexpr int kFusedBY = 8;

int fused_kz(const YeePlan* p, int gx, int gy) {
  int kz = p->fused_kz;
  if (kz <= 0) {
    if (p->fused_variant >= 3) {
      // 6 planes on the 512^3 grid (with the XCD row-run order: 1.907 ms
      // per step against 1.920 with 4 and 1.934 with 8); 4, then 2, where
      // longer chunks would leave fewer than 8192 tiles (128^3: 2 planes,
      // 0.081 against 0.086 ms per step)
      const long long tiles6 = (long long)gx * gy * ((p->d.nz + 5) / 6);
      const long long tiles4 = (long long)gx * gy * ((p->d.nz + 3) / 4);
      kz = tiles6 >= 8192 ? 6 : (tiles4 >= 8192 ? 4 : 2);
    } else {
      kz = 16;
    }
  }
  return kz > p->d.nz ? p->d.nz : kz;
}

// Variant 2 launches.  CPML: the tiles that touch no slab (in x and y with
// their x-1 / y-1 neighbours, planes p .. nz-p-1) form one box run by the
// lean SHELL == false kernel; the rest, as up to six boxes, by the SHELL
// kernel.  Both read buffer A and write disjoint cells of buffer B.
template <int MODE, bool LOSSY, int VAR, int ROWS>
void launch_v2_rows(const FusedArgs& a, const YeePlan* pl, hipStream_t st) {
  const int gx = (a.nx + 63) / 64, gy = (a.ny + ROWS - 1) / ROWS;
  const int kz = fused_kz(pl, gx, gy);
  // default: XCD row runs (3), so that the columns two x-neighbouring
  // tiles share stay in one XCD's L2 (512^3: interior reads 1.40x instead
  // of 1.94x their minimum)
  const int order = pl->xcd < 0 ? 3 : pl->xcd;
  // a tile reads the x-1 / y-1 neighbours of its cells; planes p .. nz-p-1
  // (the plane below the first is recomputed with full CPML handling)
  TileSet in, sh, shx;
  if constexpr (VAR >= 3) {
    split_tiles3(a.nx, a.ny, a.nz, MODE == kCpml ? a.p : 0, kz, 1, 0, order, in, sh, shx, ROWS,
                 0, 0, true, pl->xcd_shell);
  } else {
    split_tiles(a.nx, a.ny, a.nz, MODE == kCpml ? a.p : 0, kz, 1, 0, 0, 0, order, in, sh, ROWS);
    std::memset(&shx, 0, sizeof(shx));
  }
#define YEE_V2_LAUNCH(SH, TS)                                                                 \
  yee_fused_v2_kernel<MODE, LOSSY, SH, VAR, ROWS><<<tileset_blocks(TS), 64 * ROWS, 0, st>>>( \
      a, TS, a.fi[0], a.fi[1], a.fi[2], a.fi[3], a.fi[4], a.fi[5], a.fo[0], a.fo[1], a.fo[2], \
      a.fo[3], a.fo[4], a.fo[5])
  if (in.ntiles > 0) YEE_V2_LAUNCH(0, in);
  if constexpr (MODE == kCpml) {
    if (sh.ntiles > 0) YEE_V2_LAUNCH(1, sh);
    if constexpr (VAR >= 3)
      if (shx.ntiles > 0) YEE_V2_LAUNCH(2, shx);
  }
#undef YEE_V2_LAUNCH
}

// rows per tile: 4 unless asked otherwise
template <int MODE, bool LOSSY, int VAR>
void launch_v2(const FusedArgs& a, const YeePlan* pl, hipStream_t st) {
  if constexpr (VAR >= 3) {
    if (pl->fused_rows == 16) return launch_v2_rows<MODE, LOSSY, VAR, 16>(a, pl, st);
    if (pl->fused_rows == 8) return launch_v2_rows<MODE, LOSSY, VAR, 8>(a, pl, st);
  }
  launch_v2_rows<MODE, LOSSY, VAR, 4>(a, pl, st);
}

template <int MODE, bool LOSSY>
void launch_fused_t(const FusedArgs& a, const YeePlan* p, hipStream_t st) {
  const int gx = (a.nx + 63) / 64;
  if (p->fused_variant == 4) {
    launch_v2<MODE, LOSSY, 4>(a, p, st);
  } else if (p->fused_variant == 3) {
    launch_v2<MODE, LOSSY, 3>(a, p, st);
  } else if (p->fused_variant == 2) {
    launch_v2<MODE, LOSSY, 2>(a, p, st);
  } else if (p->fused_variant == 1) {
    const int gy = (a.ny + 3) / 4;
    const int kz = fused_kz(p, gx, gy);
    const dim3 grid(gx, gy, (a.nz + kz - 1) / kz);
    yee_fused_nolds_kernel<MODE, LOSSY><<<grid, 256, 0, st>>>(
        a, kz, a.fi[0], a.fi[1], a.fi[2], a.fi[3], a.fi[4], a.fi[5], a.fo[0], a.fo[1], a.fo[2],
        a.fo[3], a.fo[4], a.fo[5]);
  } else {
    const int gy = (a.ny + kFusedBY - 1) / kFusedBY;
    const int kz = fused_kz(p, gx, gy);
    const dim3 grid(gx, gy, (a.nz + kz - 1) / kz);
    yee_fused_kernel<MODE, LOSSY, kFusedBY><<<grid, 64 * (kFusedBY + 1), 0, st>>>(a, kz);
  }
}

int launch_fused(const YeePlan* p, float* const* fin, float* const* fout, float* pin, float* pout,
                 int n, hipStream_t st) {
  if (p->nsrc > 0 && (n < 0 || n >= p->nsteps_table))
    return fail(YEE_EINVAL, "step " + std::to_string(n) + " outside the source table (" +
                                std::to_string(p->nsteps_table) + " steps)");
  const YeeGridDesc& d = p->d;
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.p = d.boundary == YEE_BOUNDARY_CPML ? d.npml : 0;
  a.ch = d.ch;
  a.sdt = d.sponge_dt;
  a.seps0 = d.sponge_eps0;
  a.ysl = psi_slab_floats(d, 0);
  a.zsl = psi_slab_floats(d, 1);
  a.xsl = psi_slab_floats(d, 2);
  for (int c = 0; c < 6; ++c) {
    a.fi[c] = fin[c];
    a.fo[c] = fout[c];
  }
  a.cb = d.cb;
  a.ca = d.ca;
  a.mat = d.mat;
  a.pal = d.mat ? d.palette : nullptr;
  a.sponge = d.boundary == YEE_BOUNDARY_SPONGE ? d.sponge_sigma : nullptr;
  a.cpml = d.boundary == YEE_BOUNDARY_CPML ? d.cpml : nullptr;
  a.pin = pin;
  a.pout = pout;
  const bool lossy = d.ca != nullptr;
  switch (d.boundary) {
    case YEE_BOUNDARY_PERIODIC:
      lossy ? launch_fused_t<kPeriodic, true>(a, p, st) : launch_fused_t<kPeriodic, false>(a, p, st);
      break;
    case YEE_BOUNDARY_CPML:
      lossy ? launch_fused_t<kCpml, true>(a, p, st) : launch_fused_t<kCpml, false>(a, p, st);
      break;
    case YEE_BOUNDARY_SPONGE:
      lossy ? launch_fused_t<kSponge, true>(a, p, st) : launch_fused_t<kSponge, false>(a, p, st);
      break;
    default:
      lossy ? launch_fused_t<kWalls, true>(a, p, st) : launch_fused_t<kWalls, false>(a, p, st);
      break;
  }
  YEE_HIP_CHECK(hipGetLastError());
  return launch_inject(p, fout[0], fout[1], fout[2], n, 0, st);
}

}  // namespace yee

using namespace yee;

extern "C" {

int yee_step_fused(YeePlan* p, float* const* fin, float* const* fout, float* psi_in,
                   float* psi_out, int32_t n, void* stream) {
  int rc = check_fields(p, fin, psi_in);
  if (rc) return rc;
  rc = check_fields(p, fout, psi_out);
  if (rc) return rc;
  return launch_fused(p, fin, fout, psi_in, psi_out, n, (hipStream_t)stream);
}

int yee_run_fused(YeePlan* p, float* const* fa, float* const* fb, float* psi_a, float* psi_b,
                  int32_t n0, int32_t nsteps, void* stream) {
  int rc = check_fields(p, fa, psi_a);
  if (rc) return rc;
  rc = check_fields(p, fb, psi_b);
  if (rc) return rc;
  if (nsteps < 0) return fail(YEE_EINVAL, "negative step count");
  hipStream_t st = (hipStream_t)stream;
  for (int s = 0; s < nsteps; ++s) {
    const bool even = (s % 2) == 0;
    rc = launch_fused(p, even ? fa : fb, even ? fb : fa, even ? psi_a : psi_b,
                      even ? psi_b : psi_a, n0 + s, st);
    if (rc) return rc;
  }
  return YEE_OK;
}

}  // extern "C"
