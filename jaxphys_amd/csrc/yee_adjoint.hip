// jaxphys_amd: discrete adjoint (time-reversed, transposed) of the Yee step
// of yee.hip, for d(objective)/d(eps_r).  No reference counterpart: the
// reference differentiates its scan with generic jax.grad
// (src/neurosim/optimize.py:84), storing every step.  Here the exact
// transpose of one forward step is applied per step in reverse order.
//
// Forward step n (state E^n, H^{n-1/2}, psi):
//   H pass, each H cell in range:   u = (a k_a + psiA') - (b k_b + psiB'),
//     psiA' = bA psiA + cA a (slab only), a, b forward differences of E^n,
//     H' = H - ch u
//   E pass, each E cell in range:   u = (a k_a + psiA') - (b k_b + psiB'),
//     a, b backward differences of H',  E' = Ca E + Cb u;  E' *= damp;
//     E' += source (no effect on the adjoint).
// Adjoint step n, given lam = dJ/d(E^{n+1}, H^{n+1/2}) and the psi adjoints:
//   kernel A (per H cell): first the deferred psi_H-adjoint update of step
//     n+1, then lamH += transpose of the E pass (gathered from the E cells
//     that read this H value: the cell itself and its +x/+y/+z neighbours).
//   kernel B (per E cell): g = damp lamE; psi_E-adjoint update; lamE = Ca g
//     (in range) plus the transpose of the H pass gathered from the H cells
//     at the cell and its -x/-y/-z neighbours; design-box accumulation
//     acc_cb += g d with d = E^{n+1} - Ca E^n (recorded by the forward E
//     kernel), acc_ca += g E^n, from which dJ/dCb = acc_cb / Cb and
//     dJ/dCa = acc_ca.
// Reads and writes never collide: A writes only lamH and psi_H adjoints at
// its own cell and reads lamE / psi_E adjoints; B the other way round.

#include "yee_internal.h"

namespace yee {

struct AdjArgs {
  int nx, ny, nz, p;
  float ch, sdt, seps0;
  unsigned ysl, zsl, xsl;      // psi slab sizes (floats) of the y, z, x groups
  const float* cb;
  const float* ca;
  const float* sponge;         // sigma_x | sigma_y | sigma_z
  const float* cpml;           // x | y | z tables, 6 n_a floats each
  float* lpsi;                 // adjoint psi (layout of the forward psi)
  int bx0, by0, bz0, bnx, bny, bnz;  // design box (bnx == 0: none)
  const float* hist_d;         // box record of the step: E^{n+1} - Ca E^n, comp-major
  const float* hist_e;         // E^n on the box (lossy plans, for acc_ca)
  float* acc_cb;
  float* acc_ca;
  // objective terms added to the E adjoint after this step (pnx == 0: none):
  // pmap[(c * pnz + k) * pny + j) * pnx + i] = probe of E component c at
  // (px0 + i, py0 + j, pz0 + k), or -1; term = pscale * pw[q] * pvals[q]
  int px0, py0, pz0, pnx, pny, pnz;
  const int* pmap;
  const float* pw;
  const float* pvals;
  float pscale;
};

__device__ __forceinline__ unsigned adj_slot(const AdjArgs& a, int s) {
  if (s < 4) return (unsigned)s * a.ysl;
  if (s < 8) return 4 * a.ysl + (unsigned)(s - 4) * a.zsl;
  return 4 * a.ysl + 4 * a.zsl + (unsigned)(s - 8) * a.xsl;
}

// adjoint psi of slot S at (i, j, k); the cell is in the slab of S's axis
template <int S>
__device__ __forceinline__ float* adj_psi(const AdjArgs& a, int i, int j, int k) {
  const int p = a.p;
  unsigned o;
  if (S < 4) o = ((unsigned)(k * 2 * p + slab_index(j, a.ny, p))) * a.nx + i;
  else if (S < 8) o = ((unsigned)(slab_index(k, a.nz, p) * a.ny + j)) * a.nx + i;
  else o = ((unsigned)k * a.ny + j) * 2 * p + slab_index(i, a.nx, p);
  return a.lpsi + adj_slot(a, S) + o;
}

template <int AX>
__device__ __forceinline__ const float* adj_tab(const AdjArgs& a) {
  return AX == 0 ? a.cpml : (AX == 1 ? a.cpml + 6 * a.nx : a.cpml + 6 * (a.nx + a.ny));
}

template <int MODE>
__device__ __forceinline__ float adj_damp(const AdjArgs& a, int i, int j, int k) {
  if (MODE != kSponge) return 1.f;
  const float sig = ((a.sponge[i] + a.sponge[a.nx + j]) + a.sponge[a.nx + a.ny + k]) / 3.0f;
  return sig != 0.f ? expf(((-sig) * a.sdt) / a.seps0) : 1.f;
}

// The two curl terms of each component (derivative axis, psi slot):
//   Ex: a = d/dy Hz, b = d/dz Hy     Hx: a = d/dy Ez, b = d/dz Ey
//   Ey: a = d/dz Hx, b = d/dx Hz     Hy: a = d/dz Ex, b = d/dx Ez
//   Ez: a = d/dx Hy, b = d/dy Hx     Hz: a = d/dx Ey, b = d/dy Ex
template <int C> struct Terms {
  static constexpr int axa = C == 0 ? 1 : (C == 1 ? 2 : 0);
  static constexpr int axb = C == 0 ? 2 : (C == 1 ? 0 : 1);
  static constexpr int ea = C == 0 ? kExy : (C == 1 ? kEyz : kEzx);
  static constexpr int eb = C == 0 ? kExz : (C == 1 ? kEyx : kEzy);
  static constexpr int ha = C == 0 ? kHxy : (C == 1 ? kHyz : kHzx);
  static constexpr int hb = C == 0 ? kHxz : (C == 1 ? kHyx : kHzy);
};

template <int AX>
__device__ __forceinline__ int coord(int i, int j, int k) {
  return AX == 0 ? i : (AX == 1 ? j : k);
}
template <int AX>
__device__ __forceinline__ int extent(const AdjArgs& a) {
  return AX == 0 ? a.nx : (AX == 1 ? a.ny : a.nz);
}

// whether E / H component C at (i, j, k) is updated (non-periodic ranges)
template <int MODE, int C>
__device__ __forceinline__ bool e_upd(const AdjArgs& a, int i, int j, int k) {
  if (MODE == kPeriodic) return true;
  return C == 0 ? (j > 0 && k > 0) : (C == 1 ? (i > 0 && k > 0) : (i > 0 && j > 0));
}
template <int MODE, int C>
__device__ __forceinline__ bool h_upd(const AdjArgs& a, int i, int j, int k) {
  if (MODE == kPeriodic) return true;
  return C == 0 ? (j < a.ny - 1 && k < a.nz - 1)
                : (C == 1 ? (i < a.nx - 1 && k < a.nz - 1) : (i < a.nx - 1 && j < a.ny - 1));
}

// Which psi terms a launch compiles in (CPX): 0 none, 1 the y and z terms
// (j is one value per wavefront and k one per plane, so their slab tests
// are uniform branches), 2 all (x terms tested per lane).
template <int CPX, int AX>
__device__ __forceinline__ constexpr bool axis_on() {
  return AX == 0 ? CPX >= 2 : CPX >= 1;
}

// dJ/d(term a), dJ/d(term b) of E component C at the in-grid cell
// (i, j, k) from the adjoint of E^{n+1} and of psi_E^{n+1}; zero when
// `valid` is false or the cell is not updated.
template <int MODE, int CPX, int C>
__device__ __forceinline__ void eterm(const AdjArgs& a, const float* __restrict__ lamE, int i,
                                      int j, int k, bool valid, float& ga, float& gb) {
  using T = Terms<C>;
  const unsigned c = ((unsigned)k * a.ny + j) * a.nx + i;
  const float gU = a.cb[c] * (adj_damp<MODE>(a, i, j, k) * lamE[c]);
  ga = gU;
  gb = -gU;
  {
    const int p = a.p;
    const int va = coord<T::axa>(i, j, k), na = extent<T::axa>(a);
    const int vb = coord<T::axb>(i, j, k), nb = extent<T::axb>(a);
    if (axis_on<CPX, T::axa>() && (va < p || va >= na - p)) {
      const float* t = adj_tab<T::axa>(a);
      const float L = *adj_psi<T::ea>(a, i, j, k) + gU;
      ga = t[2 * na + va] * gU + t[1 * na + va] * L;
    }
    if (axis_on<CPX, T::axb>() && (vb < p || vb >= nb - p)) {
      const float* t = adj_tab<T::axb>(a);
      const float L = *adj_psi<T::eb>(a, i, j, k) - gU;
      gb = -t[2 * nb + vb] * gU + t[1 * nb + vb] * L;
    }
  }
  if (!(valid && e_upd<MODE, C>(a, i, j, k))) {
    ga = 0.f;
    gb = 0.f;
  }
}

// same for H component C from the adjoint of H^{n+1/2} and psi_H^{n+1/2}
template <int MODE, int CPX, int C>
__device__ __forceinline__ void hterm(const AdjArgs& a, const float* __restrict__ lamH, int i,
                                      int j, int k, bool valid, float& ga, float& gb) {
  using T = Terms<C>;
  const unsigned c = ((unsigned)k * a.ny + j) * a.nx + i;
  const float gU = -a.ch * lamH[c];
  ga = gU;
  gb = -gU;
  {
    const int p = a.p;
    const int va = coord<T::axa>(i, j, k), na = extent<T::axa>(a);
    const int vb = coord<T::axb>(i, j, k), nb = extent<T::axb>(a);
    if (axis_on<CPX, T::axa>() && (va < p || va >= na - p)) {
      const float* t = adj_tab<T::axa>(a);
      const float L = *adj_psi<T::ha>(a, i, j, k) + gU;
      ga = t[5 * na + va] * gU + t[4 * na + va] * L;
    }
    if (axis_on<CPX, T::axb>() && (vb < p || vb >= nb - p)) {
      const float* t = adj_tab<T::axb>(a);
      const float L = *adj_psi<T::hb>(a, i, j, k) - gU;
      gb = -t[5 * nb + vb] * gU + t[4 * nb + vb] * L;
    }
  }
  if (!(valid && h_upd<MODE, C>(a, i, j, k))) {
    ga = 0.f;
    gb = 0.f;
  }
}

// deferred psi_H-adjoint update of H component C at (i, j, k)
template <int MODE, int CPX, int C>
__device__ __forceinline__ void psi_h_defer(const AdjArgs& a, float lamh, int i, int j, int k) {
  using T = Terms<C>;
  if (!h_upd<MODE, C>(a, i, j, k)) return;
  const int p = a.p;
  const float gU = -a.ch * lamh;
  const int va = coord<T::axa>(i, j, k), na = extent<T::axa>(a);
  const int vb = coord<T::axb>(i, j, k), nb = extent<T::axb>(a);
  if (axis_on<CPX, T::axa>() && (va < p || va >= na - p)) {
    float* q = adj_psi<T::ha>(a, i, j, k);
    *q = adj_tab<T::axa>(a)[3 * na + va] * (*q + gU);
  }
  if (axis_on<CPX, T::axb>() && (vb < p || vb >= nb - p)) {
    float* q = adj_psi<T::hb>(a, i, j, k);
    *q = adj_tab<T::axb>(a)[3 * nb + vb] * (*q - gU);
  }
}

// psi_E-adjoint update of E component C at (i, j, k) (updated cells only)
template <int CPX, int C>
__device__ __forceinline__ void psi_e_update(const AdjArgs& a, float gU, int i, int j, int k) {
  using T = Terms<C>;
  const int p = a.p;
  const int va = coord<T::axa>(i, j, k), na = extent<T::axa>(a);
  const int vb = coord<T::axb>(i, j, k), nb = extent<T::axb>(a);
  if (axis_on<CPX, T::axa>() && (va < p || va >= na - p)) {
    float* q = adj_psi<T::ea>(a, i, j, k);
    *q = adj_tab<T::axa>(a)[va] * (*q + gU);
  }
  if (axis_on<CPX, T::axb>() && (vb < p || vb >= nb - p)) {
    float* q = adj_psi<T::eb>(a, i, j, k);
    *q = adj_tab<T::axb>(a)[vb] * (*q - gU);
  }
}

// neighbour index along an axis: wrapped (periodic) or clamped with a
// validity flag (other boundaries)
template <int MODE>
__device__ __forceinline__ int nb_idx(int v, int n, bool& ok) {
  if (v >= 0 && v < n) {
    ok = true;
    return v;
  }
  if (MODE == kPeriodic) {
    ok = true;
    return v < 0 ? v + n : v - n;
  }
  ok = false;
  return v < 0 ? 0 : n - 1;
}

// Column of this lane.  A paired tile holds two 32-column segments, one per
// half-wave (the low and the high x slab); every neighbour is loaded from
// memory, so the two halves need no exchange.
__device__ __forceinline__ void adj_column(const TilePos& tp, int& i, int& xl) {
  const int lane = (int)(threadIdx.x % kBX);
  const bool hi = tp.pair && lane >= 32;
  i = hi ? tp.ih0 + (lane & 31) : tp.i0 + (tp.pair ? lane & 31 : lane);
  xl = hi ? tp.ihl : tp.xl;
}

// Kernel A: per H cell, z-marching tiles.  CPX as axis_on: 0 for tiles with
// no slab within one cell, 1 for tiles whose columns are at least one cell
// away from the x slabs, 2 for any tile.
template <int MODE, int CPX>
__device__ __forceinline__ void adj_h_body(
    const AdjArgs& a, const TileSet& ts, int L, const float* __restrict__ lex,
    const float* __restrict__ ley, const float* __restrict__ lez, float* __restrict__ lhx,
    float* __restrict__ lhy, float* __restrict__ lhz) {
  TilePos tp;
  if (!tile_of_set_at(ts, L, tp)) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int k0 = tp.k0, k1 = tp.k1;
  int i, xl;
  adj_column(tp, i, xl);
  // one y row per wavefront: j is scalar, so y-slab tests are uniform
  const int j = __builtin_amdgcn_readfirstlane(tp.j0 + (int)(threadIdx.x / kBX));
  if (i >= xl || j >= tp.yl) return;
  bool vx, vy;
  const int ip = nb_idx<MODE>(i + 1, nx, vx);
  const int jp = nb_idx<MODE>(j + 1, ny, vy);
  // lane + 1 holds column i + 1 of this tile row (same half of a paired
  // tile, inside the tile, no wrap): its own terms at x+1 are taken from it
  const int lane = (int)(threadIdx.x % kBX);
  const int seg = tp.pair ? 32 : kBX;
  const bool xp_lane = (lane % seg) < seg - 1 && ip == i + 1 && i + 1 < xl;
  // the Ex and Ey terms of plane k are carried from the previous plane,
  // where they were the z+1 neighbours (k + 1 < k1 <= nz: no wrap there)
  float cxa, cxb, cya, cyb;
  eterm<MODE, CPX, 0>(a, lex, i, j, k0, true, cxa, cxb);
  eterm<MODE, CPX, 1>(a, ley, i, j, k0, true, cya, cyb);
  for (int k = k0; k < k1; ++k) {
    bool vz;
    const int kp = nb_idx<MODE>(k + 1, nz, vz);
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const float hx0 = lhx[c], hy0 = lhy[c], hz0 = lhz[c];
    if (CPX > 0) {
      psi_h_defer<MODE, CPX, 0>(a, hx0, i, j, k);
      psi_h_defer<MODE, CPX, 1>(a, hy0, i, j, k);
      psi_h_defer<MODE, CPX, 2>(a, hz0, i, j, k);
    }
    const float xa = cxa, xb = cxb, ya = cya, yb = cyb;
    float za, zb, t0, t1;
    eterm<MODE, CPX, 2>(a, lez, i, j, k, true, za, zb);
    // z+1 neighbours (also the carried terms of the next plane)
    eterm<MODE, CPX, 0>(a, lex, i, j, kp, vz, cxa, cxb);
    eterm<MODE, CPX, 1>(a, ley, i, j, kp, vz, cya, cyb);
    // x+1 neighbours: Ez's a and Ey's b at (ip, j, k)
    const float sh_za = __shfl_down(za, 1, kBX);
    const float sh_yb = __shfl_down(yb, 1, kBX);
    float zpa, ypb;
    if (xp_lane) {
      zpa = sh_za;
      ypb = sh_yb;
    } else {
      eterm<MODE, CPX, 2>(a, lez, ip, j, k, vx, zpa, t1);
      eterm<MODE, CPX, 1>(a, ley, ip, j, k, vx, t0, ypb);
    }
    // Hx: Ey's a (d/dz Hx) at c and c+z; Ez's b (d/dy Hx) at c and c+y
    float s = hx0;
    s += ya - cya;
    eterm<MODE, CPX, 2>(a, lez, i, jp, k, vy, t0, t1);
    s += zb - t1;
    lhx[c] = s;
    // Hy: Ex's b (d/dz Hy) at c and c+z; Ez's a (d/dx Hy) at c and c+x
    s = hy0;
    s += xb - cxb;
    s += za - zpa;
    lhy[c] = s;
    // Hz: Ex's a (d/dy Hz) at c and c+y; Ey's b (d/dx Hz) at c and c+x
    s = hz0;
    eterm<MODE, CPX, 0>(a, lex, i, jp, k, vy, t0, t1);
    s += xa - t0;
    s += yb - ypb;
    lhz[c] = s;
  }
}

// Kernel B: per E cell, z-marching tiles (CPX as in kernel A).
template <int MODE, bool LOSSY, int CPX>
__device__ __forceinline__ void adj_e_body(
    const AdjArgs& a, const TileSet& ts, int L, float* __restrict__ lex,
    float* __restrict__ ley, float* __restrict__ lez, const float* __restrict__ lhx,
    const float* __restrict__ lhy, const float* __restrict__ lhz) {
  TilePos tp;
  if (!tile_of_set_at(ts, L, tp)) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int k0 = tp.k0, k1 = tp.k1;
  int i, xl;
  adj_column(tp, i, xl);
  // one y row per wavefront: j is scalar, so y-slab tests are uniform
  const int j = __builtin_amdgcn_readfirstlane(tp.j0 + (int)(threadIdx.x / kBX));
  if (i >= xl || j >= tp.yl) return;
  bool vx, vy;
  const int im = nb_idx<MODE>(i - 1, nx, vx);
  const int jm = nb_idx<MODE>(j - 1, ny, vy);
  // lane - 1 holds column i - 1 of this tile row (same half of a paired
  // tile, no wrap): its own terms at x-1 are taken from it
  const int lane = (int)(threadIdx.x % kBX);
  const int seg = tp.pair ? 32 : kBX;
  const bool xm_lane = (lane % seg) > 0 && im == i - 1;
  // the Hx and Hy terms of plane k-1: computed for the first plane, then
  // carried from the previous plane's own terms (k - 1 >= k0 >= 0 there)
  float pxa, pxb, pya, pyb;
  {
    bool vz0;
    const int km0 = nb_idx<MODE>(k0 - 1, nz, vz0);
    hterm<MODE, CPX, 0>(a, lhx, i, j, km0, vz0, pxa, pxb);
    hterm<MODE, CPX, 1>(a, lhy, i, j, km0, vz0, pya, pyb);
  }
  for (int k = k0; k < k1; ++k) {
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const float d = adj_damp<MODE>(a, i, j, k);
    const float cb = a.cb[c];
    const float ca = LOSSY ? a.ca[c] : 1.f;
    const float g0 = d * lex[c], g1 = d * ley[c], g2 = d * lez[c];
    float e0 = g0, e1 = g1, e2 = g2;
    if (e_upd<MODE, 0>(a, i, j, k)) {
      if (CPX > 0) psi_e_update<CPX, 0>(a, cb * g0, i, j, k);
      e0 = LOSSY ? ca * g0 : g0;
    }
    if (e_upd<MODE, 1>(a, i, j, k)) {
      if (CPX > 0) psi_e_update<CPX, 1>(a, cb * g1, i, j, k);
      e1 = LOSSY ? ca * g1 : g1;
    }
    if (e_upd<MODE, 2>(a, i, j, k)) {
      if (CPX > 0) psi_e_update<CPX, 2>(a, cb * g2, i, j, k);
      e2 = LOSSY ? ca * g2 : g2;
    }
    float xa, xb, ya, yb, za, zb, t0, t1;
    hterm<MODE, CPX, 0>(a, lhx, i, j, k, true, xa, xb);
    hterm<MODE, CPX, 1>(a, lhy, i, j, k, true, ya, yb);
    hterm<MODE, CPX, 2>(a, lhz, i, j, k, true, za, zb);
    // x-1 neighbours: Hz's a and Hy's b at (im, j, k)
    const float sh_za = __shfl_up(za, 1, kBX);
    const float sh_yb = __shfl_up(yb, 1, kBX);
    float zma, ymb;
    if (xm_lane) {
      zma = sh_za;
      ymb = sh_yb;
    } else {
      hterm<MODE, CPX, 2>(a, lhz, im, j, k, vx, zma, t1);
      hterm<MODE, CPX, 1>(a, lhy, im, j, k, vx, t0, ymb);
    }
    // Ex: Hy's a (Ex[c+z] - Ex[c]) at c (-) and c-z (+); Hz's b at c (-), c-y (+)
    e0 += pya - ya;
    hterm<MODE, CPX, 2>(a, lhz, i, jm, k, vy, t0, t1);
    e0 += t1 - zb;
    // Ey: Hx's b (Ey[c+z] - Ey[c]) at c (-), c-z (+); Hz's a at c (-), c-x (+)
    e1 += pxb - xb;
    e1 += zma - za;
    // Ez: Hx's a (Ez[c+y] - Ez[c]) at c (-), c-y (+); Hy's b at c (-), c-x (+)
    hterm<MODE, CPX, 0>(a, lhx, i, jm, k, vy, t0, t1);
    e2 += t0 - xa;
    e2 += ymb - yb;
    pxa = xa;
    pxb = xb;
    pya = ya;
    pyb = yb;
    // design box accumulation (cells with damp == 1, all components updated)
    if (a.bnx > 0) {
      const int li = i - a.bx0, lj = j - a.by0, lk = k - a.bz0;
      if (li >= 0 && li < a.bnx && lj >= 0 && lj < a.bny && lk >= 0 && lk < a.bnz) {
        const unsigned bc = (unsigned)a.bnx * a.bny * a.bnz;
        const unsigned l = ((unsigned)lk * a.bny + lj) * a.bnx + li;
        const float gc[3] = {g0, g1, g2};
        float s1 = 0.f;
#pragma unroll
        for (int q = 0; q < 3; ++q) s1 += gc[q] * a.hist_d[q * bc + l];
        a.acc_cb[l] += s1;
        if (LOSSY && a.acc_ca) {
          float s2 = 0.f;
#pragma unroll
          for (int q = 0; q < 3; ++q) s2 += gc[q] * a.hist_e[q * bc + l];
          a.acc_ca[l] += s2;
        }
      }
    }
    lex[c] = e0;
    ley[c] = e1;
    lez[c] = e2;
  }
  // objective terms of the previous step's probes, added by the thread that
  // wrote these cells (after the march, out of its loop) exactly as the
  // probe scatter would add them
  if (a.pnx > 0) {
    const unsigned pi = (unsigned)(i - a.px0), pj = (unsigned)(j - a.py0);
    if (pi < (unsigned)a.pnx && pj < (unsigned)a.pny) {
      const unsigned pc = (unsigned)a.pnx * a.pny * a.pnz;
      const int kb = max(k0, a.pz0), ke = min(k1, a.pz0 + a.pnz);
      for (int k = kb; k < ke; ++k) {
        const unsigned c = ((unsigned)k * ny + j) * nx + i;
        const unsigned pl = ((unsigned)(k - a.pz0) * a.pny + pj) * a.pnx + pi;
        const int q0 = a.pmap[pl], q1 = a.pmap[pc + pl], q2 = a.pmap[2 * pc + pl];
        if (q0 >= 0) lex[c] += a.pscale * (a.pw ? a.pw[q0] : 1.f) * a.pvals[q0];
        if (q1 >= 0) ley[c] += a.pscale * (a.pw ? a.pw[q1] : 1.f) * a.pvals[q1];
        if (q2 >= 0) lez[c] += a.pscale * (a.pw ? a.pw[q2] : 1.f) * a.pvals[q2];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Software-pipelined bodies of the two adjoint kernels: the same terms, sums
// and order as adj_h_body / adj_e_body (so the same bits), but every load of
// plane k+1 -- the fields, and under CPML every psi value and z table row
// the plane's terms use -- is issued before the arithmetic of plane k, as
// the forward two-pass kernels do, instead of each plane (and each psi term
// inside it) waiting for its own loads.  The x and y table rows are loaded
// once per march.  eterm_q / hterm_q / psi_*_q perform the arithmetic of
// eterm / hterm / psi_h_defer / psi_e_update on the preloaded values.
// ---------------------------------------------------------------------------

// CPML data of one coordinate of one axis, loaded once and reused: whether
// it lies in the axis's slabs, its slab index (0 outside them, so that
// addresses built from it stay inside the psi arrays) and the three table
// values of the E (rows 0-2: b, c, 1/kappa) or H (rows 3-5) recursion.
struct AxAt {
  bool in;
  int s;
  float b, c, k;
};

template <int AX, int ROW>
__device__ __forceinline__ AxAt ax_at(const AdjArgs& a, int v) {
  const int n = extent<AX>(a);
  const float* t = adj_tab<AX>(a);
  AxAt r;
  r.in = v < a.p || v >= n - a.p;
  r.s = r.in ? slab_index(v, n, a.p) : 0;
  r.b = t[ROW * n + v];
  r.c = t[(ROW + 1) * n + v];
  r.k = t[(ROW + 2) * n + v];
  return r;
}

// adjoint psi of slot S at (i, j, k) from the slab index s of the slot's
// axis coordinate (adj_psi with the slab index given)
template <int S>
__device__ __forceinline__ float* adj_psi_s(const AdjArgs& a, int i, int j, int k, int s) {
  unsigned o;
  if (S < 4) o = ((unsigned)(k * 2 * a.p + s)) * a.nx + i;
  else if (S < 8) o = ((unsigned)(s * a.ny + j)) * a.nx + i;
  else o = ((unsigned)k * a.ny + j) * 2 * a.p + s;
  return a.lpsi + adj_slot(a, S) + o;
}

// the psi value of slot S (axis AX) a term at (i, j, k) needs: loaded when
// the axis is compiled in; y and z coordinates are uniform across the
// wavefront, so outside their slabs the load is skipped by a uniform
// branch, while an x coordinate outside its slab loads from the clamped
// (valid) slab index 0 and the value is not selected
template <int CPX, int AX, int S>
__device__ __forceinline__ float psi_pre(const AdjArgs& a, int i, int j, int k, const AxAt& x) {
  if (!axis_on<CPX, AX>()) return 0.f;
  if (AX != 0 && !x.in) return 0.f;
  return *adj_psi_s<S>(a, i, j, k, x.s);
}

// eterm with the psi values and table rows given: the same arithmetic
template <int MODE, int CPX, int C>
__device__ __forceinline__ void eterm_q(const AdjArgs& a, float lam, float cb, int i, int j,
                                        int k, bool valid, const AxAt& xa_, float psa,
                                        const AxAt& xb_, float psb, float& ga, float& gb) {
  using T = Terms<C>;
  const float gU = cb * (adj_damp<MODE>(a, i, j, k) * lam);
  ga = gU;
  gb = -gU;
  if (axis_on<CPX, T::axa>() && xa_.in) {
    const float L = psa + gU;
    ga = xa_.k * gU + xa_.c * L;
  }
  if (axis_on<CPX, T::axb>() && xb_.in) {
    const float L = psb - gU;
    gb = -xb_.k * gU + xb_.c * L;
  }
  if (!(valid && e_upd<MODE, C>(a, i, j, k))) {
    ga = 0.f;
    gb = 0.f;
  }
}

// hterm with the psi values and table rows given: the same arithmetic
template <int MODE, int CPX, int C>
__device__ __forceinline__ void hterm_q(const AdjArgs& a, float lamh, int i, int j, int k,
                                        bool valid, const AxAt& xa_, float psa, const AxAt& xb_,
                                        float psb, float& ga, float& gb) {
  using T = Terms<C>;
  const float gU = -a.ch * lamh;
  ga = gU;
  gb = -gU;
  if (axis_on<CPX, T::axa>() && xa_.in) {
    const float L = psa + gU;
    ga = xa_.k * gU + xa_.c * L;
  }
  if (axis_on<CPX, T::axb>() && xb_.in) {
    const float L = psb - gU;
    gb = -xb_.k * gU + xb_.c * L;
  }
  if (!(valid && h_upd<MODE, C>(a, i, j, k))) {
    ga = 0.f;
    gb = 0.f;
  }
}

// psi_h_defer with the old psi values and table rows given
template <int MODE, int CPX, int C>
__device__ __forceinline__ void psi_h_defer_q(const AdjArgs& a, float lamh, int i, int j, int k,
                                              const AxAt& xa_, float psa, const AxAt& xb_,
                                              float psb) {
  using T = Terms<C>;
  if (!h_upd<MODE, C>(a, i, j, k)) return;
  const float gU = -a.ch * lamh;
  if (axis_on<CPX, T::axa>() && xa_.in) *adj_psi_s<T::ha>(a, i, j, k, xa_.s) = xa_.b * (psa + gU);
  if (axis_on<CPX, T::axb>() && xb_.in) *adj_psi_s<T::hb>(a, i, j, k, xb_.s) = xb_.b * (psb - gU);
}

// psi_e_update with the old psi values and table rows given
template <int CPX, int C>
__device__ __forceinline__ void psi_e_update_q(const AdjArgs& a, float gU, int i, int j, int k,
                                               const AxAt& xa_, float psa, const AxAt& xb_,
                                               float psb) {
  using T = Terms<C>;
  if (axis_on<CPX, T::axa>() && xa_.in) *adj_psi_s<T::ea>(a, i, j, k, xa_.s) = xa_.b * (psa + gU);
  if (axis_on<CPX, T::axb>() && xb_.in) *adj_psi_s<T::eb>(a, i, j, k, xb_.s) = xb_.b * (psb - gU);
}

// loads of one plane of kernel A: E adjoint (and Cb) of the cell's z+1, y+1
// and (edge lanes) x+1 neighbours and of the cell itself, H adjoint of the
// cell, and under CPML every psi value the plane's terms use, with the z
// table rows of that plane
struct AdjHRaw {
  float ez, ex1, ey1, cb1, ezy, exy, cby, hx, hy, hz, ezx, eyx, cbx;
  // E psi: Ez at the cell (x, y terms), Ex / Ey at z+1, Ez / Ey at x+1
  // (edge lanes), Ez / Ex at y+1; H psi of the cell (deferred update)
  float pEzx, pEzy, pExy1, pExz1, pEyz1, pEyx1, pEzxX, pEyxX, pEzyY, pExyY;
  float pHxy, pHxz, pHyz, pHyx, pHzx, pHzy;
  AxAt zE1, zH;  // z rows: E at k+1, H at k
};

template <int MODE, int CPX>
__device__ __forceinline__ void adj_h_body_p(
    const AdjArgs& a, const TileSet& ts, int L, const float* __restrict__ lex,
    const float* __restrict__ ley, const float* __restrict__ lez, float* __restrict__ lhx,
    float* __restrict__ lhy, float* __restrict__ lhz) {
  TilePos tp;
  if (!tile_of_set_at(ts, L, tp)) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int k0 = tp.k0, k1 = tp.k1;
  int i, xl;
  adj_column(tp, i, xl);
  const int j = __builtin_amdgcn_readfirstlane(tp.j0 + (int)(threadIdx.x / kBX));
  if (i >= xl || j >= tp.yl) return;
  bool vx, vy;
  const int ip = nb_idx<MODE>(i + 1, nx, vx);
  const int jp = nb_idx<MODE>(j + 1, ny, vy);
  const int lane = (int)(threadIdx.x % kBX);
  const int seg = tp.pair ? 32 : kBX;
  const bool xp_lane = (lane % seg) < seg - 1 && ip == i + 1 && i + 1 < xl;
  // x and y table rows of the march (the column and the row are fixed)
  AxAt xE{}, xEp{}, yE{}, yEp{}, xH{}, yH{};
  if (CPX > 0) {
    xE = ax_at<0, 0>(a, i);
    xEp = ax_at<0, 0>(a, ip);
    yE = ax_at<1, 0>(a, j);
    yEp = ax_at<1, 0>(a, jp);
    xH = ax_at<0, 3>(a, i);
    yH = ax_at<1, 3>(a, j);
  }
  auto load = [&](int k) __attribute__((always_inline)) {
    AdjHRaw r;
    bool vz;
    const int kp = nb_idx<MODE>(k + 1, nz, vz);
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const unsigned cp = ((unsigned)kp * ny + j) * nx + i;
    const unsigned cy = ((unsigned)k * ny + jp) * nx + i;
    r.ez = lez[c];
    r.ex1 = lex[cp];
    r.ey1 = ley[cp];
    r.cb1 = a.cb[cp];
    r.ezy = lez[cy];
    r.exy = lex[cy];
    r.cby = a.cb[cy];
    r.hx = lhx[c];
    r.hy = lhy[c];
    r.hz = lhz[c];
    r.ezx = r.eyx = r.cbx = 0.f;
    r.pEzxX = r.pEyxX = 0.f;
    if (!xp_lane) {
      const unsigned cx = ((unsigned)k * ny + j) * nx + ip;
      r.ezx = lez[cx];
      r.eyx = ley[cx];
      r.cbx = a.cb[cx];
      if (CPX > 0) {
        r.pEzxX = psi_pre<CPX, 0, kEzx>(a, ip, j, k, xEp);
        r.pEyxX = psi_pre<CPX, 0, kEyx>(a, ip, j, k, xEp);
      }
    }
    if (CPX > 0) {
      r.zE1 = ax_at<2, 0>(a, kp);
      r.zH = ax_at<2, 3>(a, k);
      r.pEzx = psi_pre<CPX, 0, kEzx>(a, i, j, k, xE);
      r.pEzy = psi_pre<CPX, 1, kEzy>(a, i, j, k, yE);
      r.pExy1 = psi_pre<CPX, 1, kExy>(a, i, j, kp, yE);
      r.pExz1 = psi_pre<CPX, 2, kExz>(a, i, j, kp, r.zE1);
      r.pEyz1 = psi_pre<CPX, 2, kEyz>(a, i, j, kp, r.zE1);
      r.pEyx1 = psi_pre<CPX, 0, kEyx>(a, i, j, kp, xE);
      r.pEzyY = psi_pre<CPX, 1, kEzy>(a, i, jp, k, yEp);
      r.pExyY = psi_pre<CPX, 1, kExy>(a, i, jp, k, yEp);
      r.pHxy = psi_pre<CPX, 1, kHxy>(a, i, j, k, yH);
      r.pHxz = psi_pre<CPX, 2, kHxz>(a, i, j, k, r.zH);
      r.pHyz = psi_pre<CPX, 2, kHyz>(a, i, j, k, r.zH);
      r.pHyx = psi_pre<CPX, 0, kHyx>(a, i, j, k, xH);
      r.pHzx = psi_pre<CPX, 0, kHzx>(a, i, j, k, xH);
      r.pHzy = psi_pre<CPX, 1, kHzy>(a, i, j, k, yH);
    }
    return r;
  };
  float cxa, cxb, cya, cyb;
  float cbc;
  {
    const unsigned c0 = ((unsigned)k0 * ny + j) * nx + i;
    cbc = a.cb[c0];
    eterm<MODE, CPX, 0>(a, lex, i, j, k0, true, cxa, cxb);
    eterm<MODE, CPX, 1>(a, ley, i, j, k0, true, cya, cyb);
  }
  AdjHRaw cur = load(k0);
  for (int k = k0; k < k1; ++k) {
    const AdjHRaw nxt = load(k + 1 < k1 ? k + 1 : k);
    bool vz;
    const int kp = nb_idx<MODE>(k + 1, nz, vz);
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const float hx0 = cur.hx, hy0 = cur.hy, hz0 = cur.hz;
    if (CPX > 0) {
      psi_h_defer_q<MODE, CPX, 0>(a, hx0, i, j, k, yH, cur.pHxy, cur.zH, cur.pHxz);
      psi_h_defer_q<MODE, CPX, 1>(a, hy0, i, j, k, cur.zH, cur.pHyz, xH, cur.pHyx);
      psi_h_defer_q<MODE, CPX, 2>(a, hz0, i, j, k, xH, cur.pHzx, yH, cur.pHzy);
    }
    const float xa = cxa, xb = cxb, ya = cya, yb = cyb;
    float za, zb, t0, t1;
    // Ez: a along x, b along y; Ex: a along y, b along z; Ey: a along z,
    // b along x
    eterm_q<MODE, CPX, 2>(a, cur.ez, cbc, i, j, k, true, xE, cur.pEzx, yE, cur.pEzy, za, zb);
    eterm_q<MODE, CPX, 0>(a, cur.ex1, cur.cb1, i, j, kp, vz, yE, cur.pExy1, cur.zE1, cur.pExz1,
                          cxa, cxb);
    eterm_q<MODE, CPX, 1>(a, cur.ey1, cur.cb1, i, j, kp, vz, cur.zE1, cur.pEyz1, xE, cur.pEyx1,
                          cya, cyb);
    const float sh_za = __shfl_down(za, 1, kBX);
    const float sh_yb = __shfl_down(yb, 1, kBX);
    float zpa, ypb;
    if (xp_lane) {
      zpa = sh_za;
      ypb = sh_yb;
    } else {
      eterm_q<MODE, CPX, 2>(a, cur.ezx, cur.cbx, ip, j, k, vx, xEp, cur.pEzxX, yE, 0.f, zpa, t1);
      eterm_q<MODE, CPX, 1>(a, cur.eyx, cur.cbx, ip, j, k, vx, cur.zE1, 0.f, xEp, cur.pEyxX, t0,
                            ypb);
    }
    float s = hx0;
    s += ya - cya;
    eterm_q<MODE, CPX, 2>(a, cur.ezy, cur.cby, i, jp, k, vy, xE, 0.f, yEp, cur.pEzyY, t0, t1);
    s += zb - t1;
    lhx[c] = s;
    s = hy0;
    s += xb - cxb;
    s += za - zpa;
    lhy[c] = s;
    s = hz0;
    eterm_q<MODE, CPX, 0>(a, cur.exy, cur.cby, i, jp, k, vy, yEp, cur.pExyY, cur.zE1, 0.f, t0,
                          t1);
    s += xa - t0;
    s += yb - ypb;
    lhz[c] = s;
    cbc = cur.cb1;
    cur = nxt;
  }
}

// loads of one plane of kernel B: E adjoint, Cb (Ca) and H adjoint of the
// cell, H adjoint of the y-1 neighbour and (edge lanes) of the x-1
// neighbour, and under CPML every psi value the plane's terms use, with the
// z table rows of that plane
struct AdjERaw {
  float ex, ey, ez, cb, ca, hx, hy, hz, hzm, hxm, hzx, hyx;
  // H psi: the cell's Hx, Hy, Hz (both terms), Hz / Hy at x-1 (edge
  // lanes), Hz / Hx at y-1; E psi of the cell (its update)
  float pHxy, pHxz, pHyz, pHyx, pHzx, pHzy, pHzxX, pHyxX, pHzyY, pHxyY;
  float pExy, pExz, pEyz, pEyx, pEzx, pEzy;
  AxAt zH, zE;  // z rows at k: H, E
  // design box: the cell's record d
  bool inbox;
  float d0, d1, d2;
};

template <int MODE, bool LOSSY, int CPX>
__device__ __forceinline__ void adj_e_body_p(
    const AdjArgs& a, const TileSet& ts, int L, float* __restrict__ lex,
    float* __restrict__ ley, float* __restrict__ lez, const float* __restrict__ lhx,
    const float* __restrict__ lhy, const float* __restrict__ lhz) {
  TilePos tp;
  if (!tile_of_set_at(ts, L, tp)) return;
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int k0 = tp.k0, k1 = tp.k1;
  int i, xl;
  adj_column(tp, i, xl);
  const int j = __builtin_amdgcn_readfirstlane(tp.j0 + (int)(threadIdx.x / kBX));
  if (i >= xl || j >= tp.yl) return;
  bool vx, vy;
  const int im = nb_idx<MODE>(i - 1, nx, vx);
  const int jm = nb_idx<MODE>(j - 1, ny, vy);
  const int lane = (int)(threadIdx.x % kBX);
  const int seg = tp.pair ? 32 : kBX;
  const bool xm_lane = (lane % seg) > 0 && im == i - 1;
  AxAt xH{}, xHm{}, yH{}, yHm{}, xE{}, yE{};
  if (CPX > 0) {
    xH = ax_at<0, 3>(a, i);
    xHm = ax_at<0, 3>(a, im);
    yH = ax_at<1, 3>(a, j);
    yHm = ax_at<1, 3>(a, jm);
    xE = ax_at<0, 0>(a, i);
    yE = ax_at<1, 0>(a, j);
  }
  auto load = [&](int k) __attribute__((always_inline)) {
    AdjERaw r;
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const unsigned cm = ((unsigned)k * ny + jm) * nx + i;
    r.ex = lex[c];
    r.ey = ley[c];
    r.ez = lez[c];
    r.cb = a.cb[c];
    r.ca = LOSSY ? a.ca[c] : 1.f;
    r.hx = lhx[c];
    r.hy = lhy[c];
    r.hz = lhz[c];
    r.hzm = lhz[cm];
    r.hxm = lhx[cm];
    r.hzx = r.hyx = 0.f;
    r.pHzxX = r.pHyxX = 0.f;
    if (!xm_lane) {
      const unsigned cx = ((unsigned)k * ny + j) * nx + im;
      r.hzx = lhz[cx];
      r.hyx = lhy[cx];
      if (CPX > 0) {
        r.pHzxX = psi_pre<CPX, 0, kHzx>(a, im, j, k, xHm);
        r.pHyxX = psi_pre<CPX, 0, kHyx>(a, im, j, k, xHm);
      }
    }
    if (CPX > 0) {
      r.zH = ax_at<2, 3>(a, k);
      r.zE = ax_at<2, 0>(a, k);
      r.pHxy = psi_pre<CPX, 1, kHxy>(a, i, j, k, yH);
      r.pHxz = psi_pre<CPX, 2, kHxz>(a, i, j, k, r.zH);
      r.pHyz = psi_pre<CPX, 2, kHyz>(a, i, j, k, r.zH);
      r.pHyx = psi_pre<CPX, 0, kHyx>(a, i, j, k, xH);
      r.pHzx = psi_pre<CPX, 0, kHzx>(a, i, j, k, xH);
      r.pHzy = psi_pre<CPX, 1, kHzy>(a, i, j, k, yH);
      r.pHzyY = psi_pre<CPX, 1, kHzy>(a, i, jm, k, yHm);
      r.pHxyY = psi_pre<CPX, 1, kHxy>(a, i, jm, k, yHm);
      r.pExy = psi_pre<CPX, 1, kExy>(a, i, j, k, yE);
      r.pExz = psi_pre<CPX, 2, kExz>(a, i, j, k, r.zE);
      r.pEyz = psi_pre<CPX, 2, kEyz>(a, i, j, k, r.zE);
      r.pEyx = psi_pre<CPX, 0, kEyx>(a, i, j, k, xE);
      r.pEzx = psi_pre<CPX, 0, kEzx>(a, i, j, k, xE);
      r.pEzy = psi_pre<CPX, 1, kEzy>(a, i, j, k, yE);
    }
    r.inbox = false;
    if (a.bnx > 0) {
      const int li = i - a.bx0, lj = j - a.by0, lk = k - a.bz0;
      r.inbox = li >= 0 && li < a.bnx && lj >= 0 && lj < a.bny && lk >= 0 && lk < a.bnz;
      if (r.inbox) {
        const unsigned bc = (unsigned)a.bnx * a.bny * a.bnz;
        const unsigned l = ((unsigned)lk * a.bny + lj) * a.bnx + li;
        r.d0 = a.hist_d[l];
        r.d1 = a.hist_d[bc + l];
        r.d2 = a.hist_d[2 * bc + l];
      }
    }
    return r;
  };
  float pxa, pxb, pya, pyb;
  {
    bool vz0;
    const int km0 = nb_idx<MODE>(k0 - 1, nz, vz0);
    hterm<MODE, CPX, 0>(a, lhx, i, j, km0, vz0, pxa, pxb);
    hterm<MODE, CPX, 1>(a, lhy, i, j, km0, vz0, pya, pyb);
  }
  AdjERaw cur = load(k0);
  for (int k = k0; k < k1; ++k) {
    const AdjERaw nxt = load(k + 1 < k1 ? k + 1 : k);
    const unsigned c = ((unsigned)k * ny + j) * nx + i;
    const float d = adj_damp<MODE>(a, i, j, k);
    const float cb = cur.cb;
    const float ca = cur.ca;
    const float g0 = d * cur.ex, g1 = d * cur.ey, g2 = d * cur.ez;
    float e0 = g0, e1 = g1, e2 = g2;
    // Ex: a along y, b along z; Ey: a along z, b along x; Ez: a along x,
    // b along y
    if (e_upd<MODE, 0>(a, i, j, k)) {
      if (CPX > 0) psi_e_update_q<CPX, 0>(a, cb * g0, i, j, k, yE, cur.pExy, cur.zE, cur.pExz);
      e0 = LOSSY ? ca * g0 : g0;
    }
    if (e_upd<MODE, 1>(a, i, j, k)) {
      if (CPX > 0) psi_e_update_q<CPX, 1>(a, cb * g1, i, j, k, cur.zE, cur.pEyz, xE, cur.pEyx);
      e1 = LOSSY ? ca * g1 : g1;
    }
    if (e_upd<MODE, 2>(a, i, j, k)) {
      if (CPX > 0) psi_e_update_q<CPX, 2>(a, cb * g2, i, j, k, xE, cur.pEzx, yE, cur.pEzy);
      e2 = LOSSY ? ca * g2 : g2;
    }
    // Hx: a along y, b along z; Hy: a along z, b along x; Hz: a along x,
    // b along y
    float xa, xb, ya, yb, za, zb, t0, t1;
    hterm_q<MODE, CPX, 0>(a, cur.hx, i, j, k, true, yH, cur.pHxy, cur.zH, cur.pHxz, xa, xb);
    hterm_q<MODE, CPX, 1>(a, cur.hy, i, j, k, true, cur.zH, cur.pHyz, xH, cur.pHyx, ya, yb);
    hterm_q<MODE, CPX, 2>(a, cur.hz, i, j, k, true, xH, cur.pHzx, yH, cur.pHzy, za, zb);
    const float sh_za = __shfl_up(za, 1, kBX);
    const float sh_yb = __shfl_up(yb, 1, kBX);
    float zma, ymb;
    if (xm_lane) {
      zma = sh_za;
      ymb = sh_yb;
    } else {
      hterm_q<MODE, CPX, 2>(a, cur.hzx, im, j, k, vx, xHm, cur.pHzxX, yH, 0.f, zma, t1);
      hterm_q<MODE, CPX, 1>(a, cur.hyx, im, j, k, vx, cur.zH, 0.f, xHm, cur.pHyxX, t0, ymb);
    }
    e0 += pya - ya;
    hterm_q<MODE, CPX, 2>(a, cur.hzm, i, jm, k, vy, xH, 0.f, yHm, cur.pHzyY, t0, t1);
    e0 += t1 - zb;
    e1 += pxb - xb;
    e1 += zma - za;
    hterm_q<MODE, CPX, 0>(a, cur.hxm, i, jm, k, vy, yHm, cur.pHxyY, cur.zH, 0.f, t0, t1);
    e2 += t0 - xa;
    e2 += ymb - yb;
    pxa = xa;
    pxb = xb;
    pya = ya;
    pyb = yb;
    // design box accumulation (cells with damp == 1, all components
    // updated): the sums of the plain body, the record d preloaded
    if (cur.inbox) {
      const int li = i - a.bx0, lj = j - a.by0, lk = k - a.bz0;
      const unsigned l = ((unsigned)lk * a.bny + lj) * a.bnx + li;
      const float gc[3] = {g0, g1, g2};
      const float dv[3] = {cur.d0, cur.d1, cur.d2};
      float s1 = 0.f;
#pragma unroll
      for (int q = 0; q < 3; ++q) s1 += gc[q] * dv[q];
      a.acc_cb[l] += s1;
      if (LOSSY && a.acc_ca) {
        const unsigned bc = (unsigned)a.bnx * a.bny * a.bnz;
        float s2 = 0.f;
#pragma unroll
        for (int q = 0; q < 3; ++q) s2 += gc[q] * a.hist_e[q * bc + l];
        a.acc_ca[l] += s2;
      }
    }
    lex[c] = e0;
    ley[c] = e1;
    lez[c] = e2;
    cur = nxt;
  }
  if (a.pnx > 0) {
    const unsigned pi = (unsigned)(i - a.px0), pj = (unsigned)(j - a.py0);
    if (pi < (unsigned)a.pnx && pj < (unsigned)a.pny) {
      const unsigned pc = (unsigned)a.pnx * a.pny * a.pnz;
      const int kb = max(k0, a.pz0), ke = min(k1, a.pz0 + a.pnz);
      for (int k = kb; k < ke; ++k) {
        const unsigned c = ((unsigned)k * ny + j) * nx + i;
        const unsigned pl = ((unsigned)(k - a.pz0) * a.pny + pj) * a.pnx + pi;
        const int q0 = a.pmap[pl], q1 = a.pmap[pc + pl], q2 = a.pmap[2 * pc + pl];
        if (q0 >= 0) lex[c] += a.pscale * (a.pw ? a.pw[q0] : 1.f) * a.pvals[q0];
        if (q1 >= 0) ley[c] += a.pscale * (a.pw ? a.pw[q1] : 1.f) * a.pvals[q1];
        if (q2 >= 0) lez[c] += a.pscale * (a.pw ? a.pw[q2] : 1.f) * a.pvals[q2];
      }
    }
  }
}

// The adjoint kernels are latency bound (short z marches, gathers from
// neighbour rows): they are built for 8 waves per SIMD.  Of their uniform
// values the ones beyond that SGPR budget live in VGPR lanes.
#define YEE_ADJ_OCCUPANCY __attribute__((amdgpu_waves_per_eu(8)))
// the pipelined bodies hold two planes' loads, psi values included (82-89
// VGPRs): at least 5 waves per SIMD, which their registers allow without
// spills (at 6 they spill to scratch)
#define YEE_ADJ_OCCUPANCY_P __attribute__((amdgpu_waves_per_eu(5)))

// One launch per kernel and tile set (no CPML: the whole grid is one set).
template <int MODE, int CPX, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY void yee_adj_h_kernel(
    AdjArgs a, TileSet ts, const float* __restrict__ lex, const float* __restrict__ ley,
    const float* __restrict__ lez, float* __restrict__ lhx, float* __restrict__ lhy,
    float* __restrict__ lhz) {
  if (PIPE) adj_h_body_p<MODE, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
  else adj_h_body<MODE, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, bool LOSSY, int CPX, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY void yee_adj_e_kernel(
    AdjArgs a, TileSet ts, float* __restrict__ lex, float* __restrict__ ley,
    float* __restrict__ lez, const float* __restrict__ lhx, const float* __restrict__ lhy,
    const float* __restrict__ lhz) {
  if (PIPE) adj_e_body_p<MODE, LOSSY, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
  else adj_e_body<MODE, LOSSY, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, int CPX, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY_P void yee_adj_hp_kernel(
    AdjArgs a, TileSet ts, const float* __restrict__ lex, const float* __restrict__ ley,
    const float* __restrict__ lez, float* __restrict__ lhx, float* __restrict__ lhy,
    float* __restrict__ lhz) {
  if (PIPE) adj_h_body_p<MODE, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
  else adj_h_body<MODE, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, bool LOSSY, int CPX, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY_P void yee_adj_ep_kernel(
    AdjArgs a, TileSet ts, float* __restrict__ lex, float* __restrict__ ley,
    float* __restrict__ lez, const float* __restrict__ lhx, const float* __restrict__ lhy,
    const float* __restrict__ lhz) {
  if (PIPE) adj_e_body_p<MODE, LOSSY, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
  else adj_e_body<MODE, LOSSY, CPX>(a, ts, (int)blockIdx.x, lex, ley, lez, lhx, lhy, lhz);
}

// CPML: the three tile sets (slab-free interior, y/z-slab shell, x-slab
// columns) of one kernel in a single launch.  The workgroup index picks the
// set (each set's block count is a multiple of 8, so the XCD-aware order
// within every set is kept); the small, latency-bound shell tiles then run
// beside the interior tiles instead of in launches of their own.  The
// x-slab columns, the slowest tiles, take the first workgroups and the
// interior the last, so the interior's short tiles fill the launch's end.
struct AdjSets {
  TileSet in, shn, shx;
  int nb_shx, nb_shn;
};

template <int MODE, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY void yee_adj_h3_kernel(
    AdjArgs a, const AdjSets* __restrict__ setsp, const float* __restrict__ lex,
    const float* __restrict__ ley, const float* __restrict__ lez, float* __restrict__ lhx,
    float* __restrict__ lhy, float* __restrict__ lhz) {
  const AdjSets& sets = *setsp;
  int L = (int)blockIdx.x;
  if (L < sets.nb_shx) {
    if (PIPE) adj_h_body_p<MODE, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_h_body<MODE, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shx;
  if (L < sets.nb_shn) {
    if (PIPE) adj_h_body_p<MODE, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_h_body<MODE, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shn;
  if (PIPE) adj_h_body_p<MODE, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
  else adj_h_body<MODE, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, bool LOSSY, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY void yee_adj_e3_kernel(
    AdjArgs a, const AdjSets* __restrict__ setsp, float* __restrict__ lex,
    float* __restrict__ ley, float* __restrict__ lez, const float* __restrict__ lhx,
    const float* __restrict__ lhy, const float* __restrict__ lhz) {
  const AdjSets& sets = *setsp;
  int L = (int)blockIdx.x;
  if (L < sets.nb_shx) {
    if (PIPE) adj_e_body_p<MODE, LOSSY, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_e_body<MODE, LOSSY, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shx;
  if (L < sets.nb_shn) {
    if (PIPE) adj_e_body_p<MODE, LOSSY, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_e_body<MODE, LOSSY, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shn;
  if (PIPE) adj_e_body_p<MODE, LOSSY, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
  else adj_e_body<MODE, LOSSY, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY_P void yee_adj_h3p_kernel(
    AdjArgs a, const AdjSets* __restrict__ setsp, const float* __restrict__ lex,
    const float* __restrict__ ley, const float* __restrict__ lez, float* __restrict__ lhx,
    float* __restrict__ lhy, float* __restrict__ lhz) {
  const AdjSets& sets = *setsp;
  int L = (int)blockIdx.x;
  if (L < sets.nb_shx) {
    if (PIPE) adj_h_body_p<MODE, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_h_body<MODE, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shx;
  if (L < sets.nb_shn) {
    if (PIPE) adj_h_body_p<MODE, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_h_body<MODE, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shn;
  if (PIPE) adj_h_body_p<MODE, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
  else adj_h_body<MODE, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
}

template <int MODE, bool LOSSY, bool PIPE>
__global__ __launch_bounds__(kThreads) YEE_ADJ_OCCUPANCY_P void yee_adj_e3p_kernel(
    AdjArgs a, const AdjSets* __restrict__ setsp, float* __restrict__ lex,
    float* __restrict__ ley, float* __restrict__ lez, const float* __restrict__ lhx,
    const float* __restrict__ lhy, const float* __restrict__ lhz) {
  const AdjSets& sets = *setsp;
  int L = (int)blockIdx.x;
  if (L < sets.nb_shx) {
    if (PIPE) adj_e_body_p<MODE, LOSSY, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_e_body<MODE, LOSSY, 2>(a, sets.shx, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shx;
  if (L < sets.nb_shn) {
    if (PIPE) adj_e_body_p<MODE, LOSSY, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    else adj_e_body<MODE, LOSSY, 1>(a, sets.shn, L, lex, ley, lez, lhx, lhy, lhz);
    return;
  }
  L -= sets.nb_shn;
  if (PIPE) adj_e_body_p<MODE, LOSSY, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
  else adj_e_body<MODE, LOSSY, 0>(a, sets.in, L, lex, ley, lez, lhx, lhy, lhz);
}

__global__ void probe_gather_kernel(GridArgs g, const int* comp, const long long* idx, int n,
                                    float* out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) out[q] = g.f[comp[q]][idx[q]];
}

__global__ void probe_scatter_kernel(GridArgs g, const int* comp, const long long* idx,
                                     const float* w, const float* vals, float scale, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) g.f[comp[q]][idx[q]] += scale * (w ? w[q] : 1.f) * vals[q];
}

__global__ void box_gather_kernel(GridArgs g, int x0, int y0, int z0, int bnx, int bny, int bnz,
                                  float* out) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  const int lj = blockIdx.y;
  const int lk = blockIdx.z;
  if (li >= bnx) return;
  const long long bc = (long long)bnx * bny * bnz;
  const long long l = ((long long)lk * bny + lj) * bnx + li;
  const long long c = ((long long)(z0 + lk) * g.ny + (y0 + lj)) * g.nx + (x0 + li);
  out[l] = g.f[0][c];
  out[bc + l] = g.f[1][c];
  out[2 * bc + l] = g.f[2][c];
}

// The record launches of a step when the E kernel cannot record itself
// (sources injected after it, or probes without a map).  Before the step,
// record_pre copies E^n of the box into the step's record (both halves for
// lossy plans); after it, record_post turns the first half into d =
// E^{n+1} - Ca E^n in place (the E kernel's arithmetic) and gathers the
// probes: blocks [0, nb_box) work on the box (one cell per thread over the
// flattened box, three components), the rest on the probes.  Box cells fit
// 32 bits (checked on the host).
struct RecordArgs {
  int x0, y0, z0, bnx, bny, bnz;
  unsigned bc;        // box cells
  int nb_box;         // box blocks
  int lossy;          // record E^n after d
  float* box_out;
  int nprobe;
  const int* comp;
  const long long* idx;
  float* probe_out;
};

constexpr int kRecThreads = 256;

__device__ __forceinline__ long long record_cell(const GridArgs& g, const RecordArgs& r,
                                                 unsigned l) {
  const unsigned li = l % (unsigned)r.bnx;
  const unsigned row = l / (unsigned)r.bnx;
  const unsigned lj = row % (unsigned)r.bny;
  const unsigned lk = row / (unsigned)r.bny;
  return ((long long)(r.z0 + (int)lk) * g.ny + (r.y0 + (int)lj)) * g.nx + (r.x0 + (int)li);
}

__global__ __launch_bounds__(kRecThreads) void record_pre_kernel(GridArgs g, RecordArgs r) {
  const unsigned l = (unsigned)blockIdx.x * kRecThreads + threadIdx.x;
  if (l >= r.bc) return;
  const long long c = record_cell(g, r, l);
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const float e = g.f[q][c];
    r.box_out[(unsigned long long)q * r.bc + l] = e;
    if (r.lossy) r.box_out[(unsigned long long)(3 + q) * r.bc + l] = e;
  }
}

__global__ __launch_bounds__(kRecThreads) void record_post_kernel(GridArgs g, RecordArgs r) {
  const int b = blockIdx.x;
  if (b < r.nb_box) {
    const unsigned l = (unsigned)b * kRecThreads + threadIdx.x;
    if (l >= r.bc) return;
    const long long c = record_cell(g, r, l);
    float ca = 1.f;
    if (r.lossy) ca = g.mat ? g.pal[256 + g.mat[c]] : g.ca[c];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      float* o = r.box_out + (unsigned long long)q * r.bc + l;
      *o = g.f[q][c] - (r.lossy ? ca * *o : *o);
    }
    return;
  }
  const int q = (b - r.nb_box) * kRecThreads + threadIdx.x;
  if (q < r.nprobe) r.probe_out[q] = g.f[r.comp[q]][r.idx[q]];
}

int adjoint_args(const YeePlan* p, float* const* lam, float* lam_psi, const YeeDesignBox* box,
                 AdjArgs& a) {
  int rc = check_fields(p, lam, lam_psi);
  if (rc) return rc;
  const YeeGridDesc& d = p->d;
  std::memset(&a, 0, sizeof(a));
  a.nx = d.nx;
  a.ny = d.ny;
  a.nz = d.nz;
  a.p = d.boundary == YEE_BOUNDARY_CPML ? d.npml : 0;
  a.ch = d.ch;
  a.sdt = d.sponge_dt;
  a.seps0 = d.sponge_eps0;
  a.ysl = (unsigned)psi_slab_floats(d, 0);
  a.zsl = (unsigned)psi_slab_floats(d, 1);
  a.xsl = (unsigned)psi_slab_floats(d, 2);
  a.cb = d.cb;
  a.ca = d.ca;
  a.sponge = d.boundary == YEE_BOUNDARY_SPONGE ? d.sponge_sigma : nullptr;
  a.cpml = d.boundary == YEE_BOUNDARY_CPML ? d.cpml : nullptr;
  a.lpsi = lam_psi;
  if (box && box->nx > 0) {
    if (box->x0 < 0 || box->y0 < 0 || box->z0 < 0 || box->ny <= 0 || box->nz <= 0 ||
        box->x0 + box->nx > p->d.nx || box->y0 + box->ny > p->d.ny || box->z0 + box->nz > p->d.nz)
      return fail(YEE_EINVAL, "design box outside the grid");
    if (!box->d || !box->acc_cb || (d.ca && box->acc_ca && !box->e_cur))
      return fail(YEE_EINVAL, "design box needs d, acc_cb (and e_cur with acc_ca)");
    a.bx0 = box->x0;
    a.by0 = box->y0;
    a.bz0 = box->z0;
    a.bnx = box->nx;
    a.bny = box->ny;
    a.bnz = box->nz;
    a.hist_d = box->d;
    a.hist_e = box->e_cur;
    a.acc_cb = box->acc_cb;
    a.acc_ca = box->acc_ca;
  }
  return YEE_OK;
}

// Both kernels of one adjoint step, each as two launches under CPML: tiles
// with no slab within one cell (no psi code) and the slab shell.
// Device copy of the CPML launches' tile sets, kept on the plan and copied
// again only when they change (grid, tile order, planes per tile).  The
// copy is synchronous, so the launches that follow read the new sets.
const AdjSets* adj_sets_device(const YeePlan* p, const AdjSets& sets) {
  if (!p->d_adjsets) {
    void* d = nullptr;
    if (hipMalloc(&d, sizeof(AdjSets)) != hipSuccess) {
      fail(YEE_EHIP, "adjoint tile sets: device allocation failed");
      return nullptr;
    }
    p->d_adjsets = d;
    p->h_adjsets.clear();
  }
  const unsigned char* raw = reinterpret_cast<const unsigned char*>(&sets);
  if (p->h_adjsets.size() != sizeof(AdjSets) ||
      std::memcmp(p->h_adjsets.data(), raw, sizeof(AdjSets)) != 0) {
    p->h_adjsets.clear();
    const hipError_t e = hipMemcpy(p->d_adjsets, raw, sizeof(AdjSets), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      fail(YEE_EHIP, std::string("adjoint tile sets: ") + hipGetErrorString(e));
      return nullptr;
    }
    p->h_adjsets.assign(raw, raw + sizeof(AdjSets));
  }
  return static_cast<const AdjSets*>(p->d_adjsets);
}

template <int MODE, bool LOSSY, bool PIPE>
int adjoint_launch_t(const YeePlan* p, const AdjArgs& a, float* const* lam, hipStream_t st) {
  TileSet in, shn, shx;
  const int order = p->xcd < 0 ? 1 : p->xcd;
  if (a.p > 0) {
    // kernel A reads +1 and kernel B -1 neighbours: one cell of margin on
    // every side; the x-slab columns get their own launch, both slabs in
    // one paired box when each is at most 32 columns wide
    split_tiles3(a.nx, a.ny, a.nz, a.p, p->adj_kz, 1, 1, order == 2 ? 1 : order, in, shn, shx,
                 kBY, 1, 1, true);
  } else {
    split_tiles(a.nx, a.ny, a.nz, 0, p->adj_kz, 1, 1, 1, 1, order == 2 ? 1 : order, in, shn);
    std::memset(&shx, 0, sizeof(shx));
  }
  if constexpr (MODE == kCpml) {
    // block counts rounded up to multiples of 8 (the extra workgroups of a
    // plain-order set find no tile and exit)
    auto nblocks = [](const TileSet& ts) {
      if (ts.ntiles <= 0) return 0;
      const int nb = (int)tileset_blocks(ts);
      return (ts.per > 0 || ts.band > 0) ? nb : (nb + kXcds - 1) / kXcds * kXcds;
    };
    AdjSets sets;
    sets.in = in;
    sets.shn = shn;
    sets.shx = shx;
    sets.nb_shx = nblocks(shx);
    sets.nb_shn = nblocks(shn);
    const int total = sets.nb_shx + sets.nb_shn + nblocks(in);
    // the tile sets reach the kernels through a pointer: as kernel
    // arguments (about 0.9 KB) they filled the scalar registers, and the
    // kernels spent their issue slots moving spilled values in and out of
    // vector-register lanes
    const AdjSets* dsets = adj_sets_device(p, sets);
    if (!dsets) return YEE_EHIP;
    if (total > 0) {
      if constexpr (PIPE) {
        yee_adj_h3p_kernel<MODE, true><<<total, kThreads, 0, st>>>(a, dsets, lam[0], lam[1],
                                                                   lam[2], lam[3], lam[4], lam[5]);
        yee_adj_e3p_kernel<MODE, LOSSY, true><<<total, kThreads, 0, st>>>(
            a, dsets, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
      } else {
        yee_adj_h3_kernel<MODE, false><<<total, kThreads, 0, st>>>(a, dsets, lam[0], lam[1],
                                                                   lam[2], lam[3], lam[4], lam[5]);
        yee_adj_e3_kernel<MODE, LOSSY, false><<<total, kThreads, 0, st>>>(
            a, dsets, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
      }
    }
  } else {
    if (in.ntiles > 0) {
      if constexpr (PIPE) {
        yee_adj_hp_kernel<MODE, 0, true><<<tileset_blocks(in), kThreads, 0, st>>>(
            a, in, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
        yee_adj_ep_kernel<MODE, LOSSY, 0, true><<<tileset_blocks(in), kThreads, 0, st>>>(
            a, in, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
      } else {
        yee_adj_h_kernel<MODE, 0, false><<<tileset_blocks(in), kThreads, 0, st>>>(
            a, in, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
        yee_adj_e_kernel<MODE, LOSSY, 0, false><<<tileset_blocks(in), kThreads, 0, st>>>(
            a, in, lam[0], lam[1], lam[2], lam[3], lam[4], lam[5]);
      }
    }
  }
  return YEE_OK;
}

int adjoint_launch(const YeePlan* p, const AdjArgs& a, float* const* lam, hipStream_t st) {
  const bool lossy = p->d.ca != nullptr;
  int rc = YEE_OK;
#define YEE_ADJ(M)                                                       \
  if (p->adj_pipe) {                                                     \
    if (lossy) rc = adjoint_launch_t<M, true, true>(p, a, lam, st);      \
    else rc = adjoint_launch_t<M, false, true>(p, a, lam, st);           \
  } else {                                                               \
    if (lossy) rc = adjoint_launch_t<M, true, false>(p, a, lam, st);     \
    else rc = adjoint_launch_t<M, false, false>(p, a, lam, st);          \
  }
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC: YEE_ADJ(kPeriodic) break;
    case YEE_BOUNDARY_CPML: YEE_ADJ(kCpml) break;
    case YEE_BOUNDARY_SPONGE: YEE_ADJ(kSponge) break;
    default: YEE_ADJ(kWalls) break;
  }
#undef YEE_ADJ
  if (rc != YEE_OK) return rc;
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int box_check(const YeePlan* p, const int32_t* box) {
  if (!box) return YEE_OK;
  const int x0 = box[0], y0 = box[1], z0 = box[2], bnx = box[3], bny = box[4], bnz = box[5];
  if (x0 < 0 || y0 < 0 || z0 < 0 || bnx <= 0 || bny <= 0 || bnz <= 0 || x0 + bnx > p->d.nx ||
      y0 + bny > p->d.ny || z0 + bnz > p->d.nz)
    return fail(YEE_EINVAL, "box outside the grid");
  return YEE_OK;
}


}  // namespace yee

using namespace yee;

extern "C" {

int yee_adjoint_step(YeePlan* p, float* const* lam, float* lam_psi, const YeeDesignBox* box,
                     void* stream) {
  AdjArgs a;
  int rc = adjoint_args(p, lam, lam_psi, box, a);
  if (rc) return rc;
  return adjoint_launch(p, a, lam, (hipStream_t)stream);
}

int yee_probe_gather(float* const* fields, const int32_t* comp, const int64_t* idx, int32_t n,
                     float* out, void* stream) {
  if (!fields || (n > 0 && (!comp || !idx || !out)) || n < 0)
    return fail(YEE_EINVAL, "bad probe gather arguments");
  if (n == 0) return YEE_OK;
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  for (int c = 0; c < 6; ++c) g.f[c] = fields[c];
  probe_gather_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      g, comp, (const long long*)idx, n, out);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_probe_scatter_add(float* const* fields, const int32_t* comp, const int64_t* idx,
                          const float* w, const float* vals, float scale, int32_t n, void* stream) {
  if (!fields || (n > 0 && (!comp || !idx || !vals)) || n < 0)
    return fail(YEE_EINVAL, "bad probe scatter arguments");
  if (n == 0) return YEE_OK;
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  for (int c = 0; c < 6; ++c) g.f[c] = fields[c];
  probe_scatter_kernel<<<(n + 255) / 256, 256, 0, (hipStream_t)stream>>>(
      g, comp, (const long long*)idx, w, vals, scale, n);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_box_gather(float* const* fields, int32_t nx, int32_t ny, int32_t nz, const int32_t* box,
                   float* out, void* stream) {
  if (!fields || !box || !out) return fail(YEE_EINVAL, "bad box gather arguments");
  const int x0 = box[0], y0 = box[1], z0 = box[2], bnx = box[3], bny = box[4], bnz = box[5];
  if (x0 < 0 || y0 < 0 || z0 < 0 || bnx <= 0 || bny <= 0 || bnz <= 0 || x0 + bnx > nx ||
      y0 + bny > ny || z0 + bnz > nz)
    return fail(YEE_EINVAL, "box outside the grid");
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  for (int c = 0; c < 3; ++c) g.f[c] = fields[c];
  const dim3 grid((bnx + 63) / 64, bny, bnz);
  box_gather_kernel<<<grid, 64, 0, (hipStream_t)stream>>>(g, x0, y0, z0, bnx, bny, bnz, out);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_run_record(YeePlan* p, float* const* fields, float* psi, int32_t n0, int32_t nsteps,
                   int32_t nprobe, const int32_t* probe_comp, const int64_t* probe_idx,
                   float* probe_out, const int32_t* box, float* box_out,
                   const int32_t* probe_box, const int32_t* probe_map, void* stream) {
  int rc = check_fields(p, fields, psi);
  if (rc) return rc;
  if (nsteps < 0 || nprobe < 0) return fail(YEE_EINVAL, "negative count");
  if (nprobe > 0 && (!probe_comp || !probe_idx || !probe_out))
    return fail(YEE_EINVAL, "probe arrays missing");
  rc = box_check(p, box);
  if (rc) return rc;
  if (box && !box_out) return fail(YEE_EINVAL, "box output missing");
  const bool mapped = nprobe > 0 && probe_box;
  if (mapped) {
    rc = box_check(p, probe_box);
    if (rc) return rc;
    if (!probe_map) return fail(YEE_EINVAL, "probe map missing");
    if (3LL * probe_box[3] * probe_box[4] * probe_box[5] >= (1LL << 31))
      return fail(YEE_EINVAL, "probe map box too large");
  }
  const GridArgs g = make_args(p, fields, psi);
  hipStream_t st = (hipStream_t)stream;
  const long long bc3 = box ? 3LL * box[3] * box[4] * box[5] : 0;
  const bool lossy = p->d.ca != nullptr;
  const long long slot = lossy ? 2 * bc3 : bc3;  // floats of one step's box record
  if (slot >= (1LL << 32)) return fail(YEE_EINVAL, "design box too large to record");
  // inside the E kernel when no source is injected after it and the probes
  // (if any) come with a map; else record launches before and after the step
  const bool in_kernel = (box || nprobe > 0) && p->nsrc <= YEE_FUSED_SOURCES &&
                         (nprobe == 0 || mapped);
  RecArgs e;
  std::memset(&e, 0, sizeof(e));
  RecordArgs r;
  std::memset(&r, 0, sizeof(r));
  if (box) {
    e.bx0 = r.x0 = box[0];
    e.by0 = r.y0 = box[1];
    e.bz0 = r.z0 = box[2];
    e.bnx = r.bnx = box[3];
    e.bny = r.bny = box[4];
    e.bnz = r.bnz = box[5];
    e.bc = r.bc = (unsigned)(bc3 / 3);
    r.nb_box = (int)((r.bc + kRecThreads - 1) / kRecThreads);
    r.lossy = lossy ? 1 : 0;
  }
  if (mapped) {
    e.pmap = probe_map;
    e.px0 = probe_box[0]; e.py0 = probe_box[1]; e.pz0 = probe_box[2];
    e.pnx = probe_box[3]; e.pny = probe_box[4]; e.pnz = probe_box[5];
  }
  r.nprobe = nprobe;
  r.comp = probe_comp;
  r.idx = (const long long*)probe_idx;
  const int nblocks = r.nb_box + (nprobe + kRecThreads - 1) / kRecThreads;
  for (int s = 0; s < nsteps; ++s) {
    float* bo = box ? box_out + s * slot : nullptr;
    float* po = nprobe > 0 ? probe_out + (long long)s * nprobe : nullptr;
    if (!in_kernel && box) {
      r.box_out = bo;
      record_pre_kernel<<<r.nb_box, kRecThreads, 0, st>>>(g, r);
      YEE_HIP_CHECK(hipGetLastError());
    }
    rc = launch_h(p, g, st);
    if (rc) return rc;
    if (in_kernel) {
      e.box_out = bo;
      e.probe_out = po;
      rc = launch_e(p, g, n0 + s, st, &e);
      if (rc) return rc;
      continue;
    }
    rc = launch_e(p, g, n0 + s, st);
    if (rc) return rc;
    if (nblocks > 0) {
      r.box_out = bo;
      r.probe_out = po;
      record_post_kernel<<<nblocks, kRecThreads, 0, st>>>(g, r);
      YEE_HIP_CHECK(hipGetLastError());
    }
  }
  return YEE_OK;
}

int yee_adjoint_run(YeePlan* p, float* const* lam, float* lam_psi, int32_t n_lo, int32_t n_hi,
                    int32_t nprobe, const int32_t* probe_comp, const int64_t* probe_idx,
                    const float* w, const float* probe_hist, float scale, const int32_t* box,
                    const float* box_hist, float* acc_cb, float* acc_ca,
                    const int32_t* probe_box, const int32_t* probe_map, void* stream) {
  if (n_lo < 0 || n_hi < n_lo) return fail(YEE_EINVAL, "bad step range");
  if (nprobe > 0 && (!probe_comp || !probe_idx || !probe_hist))
    return fail(YEE_EINVAL, "probe arrays missing");
  int rc = box_check(p, box);
  if (rc) return rc;
  const bool mapped = nprobe > 0 && probe_box;
  if (mapped) {
    rc = box_check(p, probe_box);
    if (rc) return rc;
    if (!probe_map) return fail(YEE_EINVAL, "probe map missing");
    if (3LL * probe_box[3] * probe_box[4] * probe_box[5] >= (1LL << 31))
      return fail(YEE_EINVAL, "probe map box too large");
  }
  YeeDesignBox b;
  std::memset(&b, 0, sizeof(b));
  long long bc3 = 0;
  if (box) {
    if (!box_hist || !acc_cb) return fail(YEE_EINVAL, "box history / accumulator missing");
    b.x0 = box[0]; b.y0 = box[1]; b.z0 = box[2]; b.nx = box[3]; b.ny = box[4]; b.nz = box[5];
    b.acc_cb = acc_cb;
    b.acc_ca = acc_ca;
    bc3 = 3LL * box[3] * box[4] * box[5];
  }
  // one step's box record: d (and E^n for lossy plans), as yee_run_record
  const long long slot = p->d.ca ? 2 * bc3 : bc3;
  hipStream_t st = (hipStream_t)stream;
  GridArgs gl;
  std::memset(&gl, 0, sizeof(gl));
  for (int c = 0; c < 6; ++c) gl.f[c] = lam[c];
  for (int n = n_hi - 1; n >= n_lo; --n) {
    if (box) {
      b.d = box_hist + (long long)(n - n_lo) * slot;
      b.e_cur = p->d.ca ? b.d + bc3 : nullptr;
    }
    AdjArgs a;
    rc = adjoint_args(p, lam, lam_psi, box ? &b : nullptr, a);
    if (rc) return rc;
    if (mapped && n >= 1) {
      a.px0 = probe_box[0]; a.py0 = probe_box[1]; a.pz0 = probe_box[2];
      a.pnx = probe_box[3]; a.pny = probe_box[4]; a.pnz = probe_box[5];
      a.pmap = probe_map;
      a.pw = w;
      a.pvals = probe_hist + (long long)(n - 1) * nprobe;
      a.pscale = scale;
    }
    rc = adjoint_launch(p, a, lam, st);
    if (rc) return rc;
    if (!mapped && n >= 1 && nprobe > 0) {
      probe_scatter_kernel<<<(nprobe + 255) / 256, 256, 0, st>>>(
          gl, probe_comp, (const long long*)probe_idx, w,
          probe_hist + (long long)(n - 1) * nprobe, scale, nprobe);
      YEE_HIP_CHECK(hipGetLastError());
    }
  }
  return YEE_OK;
}

}  // extern "C"
