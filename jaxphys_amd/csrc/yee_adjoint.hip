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
//     acc_cb += g (E^{n+1} - Ca E^n), acc_ca += g E^n, from which
//     dJ/dCb = acc_cb / Cb and dJ/dCa = acc_ca.
// Reads and writes never collide: A writes only lamH and psi_H adjoints at
// its own cell and reads lamE / psi_E adjoints; B the other way round.

#include "yee_internal.h"

namespace yee {

struct AdjArgs {
  GridArgs g;                  // forward coefficients (f[] unused)
  float* lam[6];               // adjoint fields, in/out
  float* lpsi[kNumPsi];        // adjoint psi (same layout as forward psi)
  int bx0, by0, bz0, bnx, bny, bnz;  // design box (bnx == 0: none)
  const float* e_next;         // box history E^{n+1}, comp-major
  const float* e_cur;          // box history E^n
  float* acc_cb;
  float* acc_ca;
};

template <int MODE>
struct Geo {
  const AdjArgs& a;
  int nx, ny, nz, p;
  unsigned P;
  __device__ Geo(const AdjArgs& a_) : a(a_), nx(a_.g.nx), ny(a_.g.ny), nz(a_.g.nz), p(a_.g.p) {
    P = (unsigned)nx * ny;
  }
  __device__ unsigned idx(int i, int j, int k) const { return ((unsigned)k * ny + j) * nx + i; }
  // wrap (periodic) or validity (others) of a neighbour index
  __device__ bool fix(int& v, int n) const {
    if (v >= 0 && v < n) return true;
    if (MODE != kPeriodic) return false;
    v = v < 0 ? v + n : v - n;
    return true;
  }
  __device__ bool e_in(int comp, int i, int j, int k) const {
    if (MODE == kPeriodic) return true;
    if (comp == 0) return j > 0 && k > 0;
    if (comp == 1) return i > 0 && k > 0;
    return i > 0 && j > 0;
  }
  __device__ bool h_in(int comp, int i, int j, int k) const {
    if (MODE == kPeriodic) return true;
    if (comp == 0) return j < ny - 1 && k < nz - 1;
    if (comp == 1) return i < nx - 1 && k < nz - 1;
    return i < nx - 1 && j < ny - 1;
  }
  __device__ float damp(int i, int j, int k) const {
    if (MODE != kSponge) return 1.f;
    const float sig = ((a.g.ssx[i] + a.g.ssy[j]) + a.g.ssz[k]) / 3.0f;
    return sig != 0.f ? expf(((-sig) * a.g.sdt) / a.g.seps0) : 1.f;
  }
  __device__ bool slab(int v, int n) const { return MODE == kCpml && (v < p || v >= n - p); }
  // psi slot address for derivative axis ax (0 x, 1 y, 2 z) at (i, j, k)
  __device__ float* psi_at(float* base, int ax, int i, int j, int k) const {
    if (ax == 1) return base + ((unsigned)(k * 2 * p + slab_index(j, ny, p)) * nx + i);
    if (ax == 2) return base + ((unsigned)(slab_index(k, nz, p) * ny + j) * nx + i);
    return base + ((unsigned)k * ny + j) * 2 * p + slab_index(i, nx, p);
  }
  __device__ const float* tab(int ax) const { return ax == 0 ? a.g.cpx : (ax == 1 ? a.g.cpy : a.g.cpz); }
  __device__ int nax(int ax) const { return ax == 0 ? nx : (ax == 1 ? ny : nz); }
};

// The two curl terms of each component: (derivative axis of term a, psi slot
// of a, derivative axis of term b, psi slot of b), for E and for H.
//   Ex: a = d/dy Hz, b = d/dz Hy     Hx: a = d/dy Ez, b = d/dz Ey
//   Ey: a = d/dz Hx, b = d/dx Hz     Hy: a = d/dz Ex, b = d/dx Ez
//   Ez: a = d/dx Hy, b = d/dy Hx     Hz: a = d/dx Ey, b = d/dy Ex
__device__ __constant__ int kAxA[3] = {1, 2, 0};
__device__ __constant__ int kAxB[3] = {2, 0, 1};
__device__ __constant__ int kPsiEA[3] = {kExy, kEyz, kEzx};
__device__ __constant__ int kPsiEB[3] = {kExz, kEyx, kEzy};
__device__ __constant__ int kPsiHA[3] = {kHxy, kHyz, kHzx};
__device__ __constant__ int kPsiHB[3] = {kHxz, kHyx, kHzy};

// d J / d a  and  d J / d b  of E component `comp` at (i, j, k), from the
// adjoint of E^{n+1} (lamE) and of psi_E^{n+1}.  Zero out of range.
template <int MODE>
__device__ void e_term_adj(const Geo<MODE>& G, int comp, int i, int j, int k, float& ga, float& gb) {
  ga = 0.f;
  gb = 0.f;
  if (!G.e_in(comp, i, j, k)) return;
  const unsigned c = G.idx(i, j, k);
  const float gU = G.a.g.cb[c] * (G.damp(i, j, k) * G.a.lam[comp][c]);
  ga = gU;
  gb = -gU;
  const int axa = kAxA[comp], axb = kAxB[comp];
  const int va = axa == 0 ? i : (axa == 1 ? j : k);
  const int vb = axb == 0 ? i : (axb == 1 ? j : k);
  if (G.slab(va, G.nax(axa))) {
    const float* t = G.tab(axa);
    const int n = G.nax(axa);
    const float L = *G.psi_at(G.a.lpsi[kPsiEA[comp]], axa, i, j, k) + gU;
    ga = t[2 * n + va] * gU + t[1 * n + va] * L;
  }
  if (G.slab(vb, G.nax(axb))) {
    const float* t = G.tab(axb);
    const int n = G.nax(axb);
    const float L = *G.psi_at(G.a.lpsi[kPsiEB[comp]], axb, i, j, k) - gU;
    gb = -t[2 * n + vb] * gU + t[1 * n + vb] * L;
  }
}

// same for H component `comp`: uses lamH (already the total adjoint of
// H^{n+1/2}) and the adjoint of psi_H^{n+1/2}.
template <int MODE>
__device__ void h_term_adj(const Geo<MODE>& G, int comp, int i, int j, int k, float& ga, float& gb) {
  ga = 0.f;
  gb = 0.f;
  if (!G.h_in(comp, i, j, k)) return;
  const unsigned c = G.idx(i, j, k);
  const float gU = -G.a.g.ch * G.a.lam[3 + comp][c];
  ga = gU;
  gb = -gU;
  const int axa = kAxA[comp], axb = kAxB[comp];
  const int va = axa == 0 ? i : (axa == 1 ? j : k);
  const int vb = axb == 0 ? i : (axb == 1 ? j : k);
  if (G.slab(va, G.nax(axa))) {
    const float* t = G.tab(axa);
    const int n = G.nax(axa);
    const float L = *G.psi_at(G.a.lpsi[kPsiHA[comp]], axa, i, j, k) + gU;
    ga = t[5 * n + va] * gU + t[4 * n + va] * L;
  }
  if (G.slab(vb, G.nax(axb))) {
    const float* t = G.tab(axb);
    const int n = G.nax(axb);
    const float L = *G.psi_at(G.a.lpsi[kPsiHB[comp]], axb, i, j, k) - gU;
    gb = -t[5 * n + vb] * gU + t[4 * n + vb] * L;
  }
}

template <int MODE>
__device__ void e_term_adj_at(const Geo<MODE>& G, int comp, int i, int j, int k, float& ga, float& gb) {
  ga = 0.f;
  gb = 0.f;
  if (!G.fix(i, G.nx) || !G.fix(j, G.ny) || !G.fix(k, G.nz)) return;
  e_term_adj<MODE>(G, comp, i, j, k, ga, gb);
}

template <int MODE>
__device__ void h_term_adj_at(const Geo<MODE>& G, int comp, int i, int j, int k, float& ga, float& gb) {
  ga = 0.f;
  gb = 0.f;
  if (!G.fix(i, G.nx) || !G.fix(j, G.ny) || !G.fix(k, G.nz)) return;
  h_term_adj<MODE>(G, comp, i, j, k, ga, gb);
}

// Kernel A: per H cell.
template <int MODE>
__global__ __launch_bounds__(kThreads) void yee_adj_h_kernel(AdjArgs a) {
  const int i = blockIdx.x * kBX + (threadIdx.x % kBX);
  const int j = blockIdx.y * kBY + (threadIdx.x / kBX);
  const int k = blockIdx.z;
  Geo<MODE> G(a);
  if (i >= G.nx || j >= G.ny) return;
  const unsigned c = G.idx(i, j, k);
  const float ch = a.g.ch;
  // deferred psi_H-adjoint update of the step after this one
  if (MODE == kCpml) {
    for (int comp = 0; comp < 3; ++comp) {
      if (!G.h_in(comp, i, j, k)) continue;
      const float gU = -ch * a.lam[3 + comp][c];
      const int axa = kAxA[comp], axb = kAxB[comp];
      const int va = axa == 0 ? i : (axa == 1 ? j : k);
      const int vb = axb == 0 ? i : (axb == 1 ? j : k);
      if (G.slab(va, G.nax(axa))) {
        float* q = G.psi_at(a.lpsi[kPsiHA[comp]], axa, i, j, k);
        *q = G.tab(axa)[3 * G.nax(axa) + va] * (*q + gU);
      }
      if (G.slab(vb, G.nax(axb))) {
        float* q = G.psi_at(a.lpsi[kPsiHB[comp]], axb, i, j, k);
        *q = G.tab(axb)[3 * G.nax(axb) + vb] * (*q - gU);
      }
    }
  }
  float ga0, gb0, ga1, gb1;
  // Hx: a_y (Ey's z term, Hx[c] - Hx[c-z]) at c and c+z; b_z (Ez's y term) at c and c+y
  {
    float s = a.lam[3][c];
    e_term_adj_at<MODE>(G, 1, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 1, i, j, k + 1, ga1, gb1);
    s += ga0 - ga1;
    e_term_adj_at<MODE>(G, 2, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 2, i, j + 1, k, ga1, gb1);
    s += gb0 - gb1;
    a.lam[3][c] = s;
  }
  // Hy: b_x (Ex's z term) at c and c+z; a_z (Ez's x term) at c and c+x
  {
    float s = a.lam[4][c];
    e_term_adj_at<MODE>(G, 0, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 0, i, j, k + 1, ga1, gb1);
    s += gb0 - gb1;
    e_term_adj_at<MODE>(G, 2, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 2, i + 1, j, k, ga1, gb1);
    s += ga0 - ga1;
    a.lam[4][c] = s;
  }
  // Hz: a_x (Ex's y term) at c and c+y; b_y (Ey's x term) at c and c+x
  {
    float s = a.lam[5][c];
    e_term_adj_at<MODE>(G, 0, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 0, i, j + 1, k, ga1, gb1);
    s += ga0 - ga1;
    e_term_adj_at<MODE>(G, 1, i, j, k, ga0, gb0);
    e_term_adj_at<MODE>(G, 1, i + 1, j, k, ga1, gb1);
    s += gb0 - gb1;
    a.lam[5][c] = s;
  }
}

// Kernel B: per E cell.
template <int MODE, bool LOSSY>
__global__ __launch_bounds__(kThreads) void yee_adj_e_kernel(AdjArgs a) {
  const int i = blockIdx.x * kBX + (threadIdx.x % kBX);
  const int j = blockIdx.y * kBY + (threadIdx.x / kBX);
  const int k = blockIdx.z;
  Geo<MODE> G(a);
  if (i >= G.nx || j >= G.ny) return;
  const unsigned c = G.idx(i, j, k);
  const float d = G.damp(i, j, k);
  const float cb = a.g.cb[c];
  const float ca = LOSSY ? a.g.ca[c] : 1.f;
  float ga0, gb0, ga1, gb1;
  float newE[3];
  float gcomp[3];
  for (int comp = 0; comp < 3; ++comp) {
    const float g = d * a.lam[comp][c];
    gcomp[comp] = g;
    if (G.e_in(comp, i, j, k)) {
      const float gU = cb * g;
      if (MODE == kCpml) {
        const int axa = kAxA[comp], axb = kAxB[comp];
        const int va = axa == 0 ? i : (axa == 1 ? j : k);
        const int vb = axb == 0 ? i : (axb == 1 ? j : k);
        if (G.slab(va, G.nax(axa))) {
          float* q = G.psi_at(a.lpsi[kPsiEA[comp]], axa, i, j, k);
          *q = G.tab(axa)[va] * (*q + gU);
        }
        if (G.slab(vb, G.nax(axb))) {
          float* q = G.psi_at(a.lpsi[kPsiEB[comp]], axb, i, j, k);
          *q = G.tab(axb)[vb] * (*q - gU);
        }
      }
      newE[comp] = LOSSY ? ca * g : g;
    } else {
      newE[comp] = g;
    }
  }
  // Ex: Hy's a (Ex[c+z]-Ex[c]) at c (-) and c-z (+); Hz's b (Ex[c+y]-Ex[c]) at c (-), c-y (+)
  h_term_adj_at<MODE>(G, 1, i, j, k, ga0, gb0);
  h_term_adj_at<MODE>(G, 1, i, j, k - 1, ga1, gb1);
  newE[0] += ga1 - ga0;
  const float hy_b0 = gb0;  // Hy's b at c, reused for Ez
  h_term_adj_at<MODE>(G, 2, i, j, k, ga0, gb0);
  h_term_adj_at<MODE>(G, 2, i, j - 1, k, ga1, gb1);
  newE[0] += gb1 - gb0;
  const float hz_a0 = ga0;  // Hz's a at c, reused for Ey
  // Ey: Hx's b (Ey[c+z]-Ey[c]) at c (-), c-z (+); Hz's a (Ey[c+x]-Ey[c]) at c (-), c-x (+)
  h_term_adj_at<MODE>(G, 0, i, j, k, ga0, gb0);
  const float hx_a0 = ga0;  // Hx's a at c, reused for Ez
  h_term_adj_at<MODE>(G, 0, i, j, k - 1, ga1, gb1);
  newE[1] += gb1 - gb0;
  h_term_adj_at<MODE>(G, 2, i - 1, j, k, ga1, gb1);
  newE[1] += ga1 - hz_a0;
  // Ez: Hx's a (Ez[c+y]-Ez[c]) at c (-), c-y (+); Hy's b (Ez[c+x]-Ez[c]) at c (-), c-x (+)
  h_term_adj_at<MODE>(G, 0, i, j - 1, k, ga1, gb1);
  newE[2] += ga1 - hx_a0;
  h_term_adj_at<MODE>(G, 1, i - 1, j, k, ga1, gb1);
  newE[2] += gb1 - hy_b0;
  // design box accumulation (cells with damp == 1, all components in range)
  if (a.bnx > 0) {
    const int li = i - a.bx0, lj = j - a.by0, lk = k - a.bz0;
    if (li >= 0 && li < a.bnx && lj >= 0 && lj < a.bny && lk >= 0 && lk < a.bnz) {
      const long long bc = (unsigned)a.bnx * a.bny * a.bnz;
      const long long l = ((unsigned)lk * a.bny + lj) * a.bnx + li;
      float s1 = 0.f, s2 = 0.f;
      for (int comp = 0; comp < 3; ++comp) {
        const float en = a.e_next[comp * bc + l], ec = a.e_cur[comp * bc + l];
        s1 += gcomp[comp] * (en - (LOSSY ? ca * ec : ec));
        s2 += gcomp[comp] * ec;
      }
      a.acc_cb[l] += s1;
      if (LOSSY && a.acc_ca) a.acc_ca[l] += s2;
    }
  }
  a.lam[0][c] = newE[0];
  a.lam[1][c] = newE[1];
  a.lam[2][c] = newE[2];
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


int adjoint_args(const YeePlan* p, float* const* lam, float* lam_psi, const YeeDesignBox* box,
                 AdjArgs& a) {
  int rc = check_fields(p, lam, lam_psi);
  if (rc) return rc;
  std::memset(&a, 0, sizeof(a));
  a.g = make_args(p, lam, lam_psi);
  for (int c = 0; c < 6; ++c) a.lam[c] = lam[c];
  for (int s = 0; s < kNumPsi; ++s) a.lpsi[s] = a.g.psi[s];
  if (box && box->nx > 0) {
    if (box->x0 < 0 || box->y0 < 0 || box->z0 < 0 || box->ny <= 0 || box->nz <= 0 ||
        box->x0 + box->nx > p->d.nx || box->y0 + box->ny > p->d.ny || box->z0 + box->nz > p->d.nz)
      return fail(YEE_EINVAL, "design box outside the grid");
    if (!box->e_next || !box->e_cur || !box->acc_cb)
      return fail(YEE_EINVAL, "design box needs e_next, e_cur and acc_cb");
    a.bx0 = box->x0;
    a.by0 = box->y0;
    a.bz0 = box->z0;
    a.bnx = box->nx;
    a.bny = box->ny;
    a.bnz = box->nz;
    a.e_next = box->e_next;
    a.e_cur = box->e_cur;
    a.acc_cb = box->acc_cb;
    a.acc_ca = box->acc_ca;
  }
  return YEE_OK;
}

int adjoint_launch(const YeePlan* p, const AdjArgs& a, hipStream_t st) {
  const dim3 grid((p->d.nx + kBX - 1) / kBX, (p->d.ny + kBY - 1) / kBY, p->d.nz);
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC: yee_adj_h_kernel<kPeriodic><<<grid, kThreads, 0, st>>>(a); break;
    case YEE_BOUNDARY_CPML: yee_adj_h_kernel<kCpml><<<grid, kThreads, 0, st>>>(a); break;
    case YEE_BOUNDARY_SPONGE: yee_adj_h_kernel<kSponge><<<grid, kThreads, 0, st>>>(a); break;
    default: yee_adj_h_kernel<kWalls><<<grid, kThreads, 0, st>>>(a); break;
  }
  YEE_HIP_CHECK(hipGetLastError());
  const bool lossy = p->d.ca != nullptr;
#define YEE_ADJ_E(M)                                                         \
  if (lossy) yee_adj_e_kernel<M, true><<<grid, kThreads, 0, st>>>(a);       \
  else yee_adj_e_kernel<M, false><<<grid, kThreads, 0, st>>>(a);
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC: YEE_ADJ_E(kPeriodic) break;
    case YEE_BOUNDARY_CPML: YEE_ADJ_E(kCpml) break;
    case YEE_BOUNDARY_SPONGE: YEE_ADJ_E(kSponge) break;
    default: YEE_ADJ_E(kWalls) break;
  }
#undef YEE_ADJ_E
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

void box_gather_launch(const GridArgs& g, const int32_t* box, float* out, hipStream_t st) {
  const dim3 grid((box[3] + 63) / 64, box[4], box[5]);
  box_gather_kernel<<<grid, 64, 0, st>>>(g, box[0], box[1], box[2], box[3], box[4], box[5], out);
}

}  // namespace yee

using namespace yee;

extern "C" {

int yee_adjoint_step(YeePlan* p, float* const* lam, float* lam_psi, const YeeDesignBox* box,
                     void* stream) {
  AdjArgs a;
  int rc = adjoint_args(p, lam, lam_psi, box, a);
  if (rc) return rc;
  return adjoint_launch(p, a, (hipStream_t)stream);
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
                   float* probe_out, const int32_t* box, float* box_out, void* stream) {
  int rc = check_fields(p, fields, psi);
  if (rc) return rc;
  if (nsteps < 0 || nprobe < 0) return fail(YEE_EINVAL, "negative count");
  if (nprobe > 0 && (!probe_comp || !probe_idx || !probe_out))
    return fail(YEE_EINVAL, "probe arrays missing");
  rc = box_check(p, box);
  if (rc) return rc;
  if (box && !box_out) return fail(YEE_EINVAL, "box output missing");
  const GridArgs g = make_args(p, fields, psi);
  hipStream_t st = (hipStream_t)stream;
  const long long bc3 = box ? 3LL * box[3] * box[4] * box[5] : 0;
  for (int s = 0; s < nsteps; ++s) {
    rc = launch_h(p, g, st);
    if (rc) return rc;
    rc = launch_e(p, g, n0 + s, st);
    if (rc) return rc;
    if (nprobe > 0) {
      probe_gather_kernel<<<(nprobe + 255) / 256, 256, 0, st>>>(
          g, probe_comp, (const long long*)probe_idx, nprobe, probe_out + (long long)s * nprobe);
      YEE_HIP_CHECK(hipGetLastError());
    }
    if (box) {
      box_gather_launch(g, box, box_out + s * bc3, st);
      YEE_HIP_CHECK(hipGetLastError());
    }
  }
  return YEE_OK;
}

int yee_adjoint_run(YeePlan* p, float* const* lam, float* lam_psi, int32_t n_lo, int32_t n_hi,
                    int32_t nprobe, const int32_t* probe_comp, const int64_t* probe_idx,
                    const float* w, const float* probe_hist, float scale, const int32_t* box,
                    const float* box_hist, float* acc_cb, float* acc_ca, void* stream) {
  if (n_lo < 0 || n_hi < n_lo) return fail(YEE_EINVAL, "bad step range");
  if (nprobe > 0 && (!probe_comp || !probe_idx || !probe_hist))
    return fail(YEE_EINVAL, "probe arrays missing");
  int rc = box_check(p, box);
  if (rc) return rc;
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
  hipStream_t st = (hipStream_t)stream;
  GridArgs gl;
  std::memset(&gl, 0, sizeof(gl));
  for (int c = 0; c < 6; ++c) gl.f[c] = lam[c];
  for (int n = n_hi - 1; n >= n_lo; --n) {
    if (box) {
      b.e_next = box_hist + (long long)(n - n_lo + 1) * bc3;
      b.e_cur = box_hist + (long long)(n - n_lo) * bc3;
    }
    AdjArgs a;
    rc = adjoint_args(p, lam, lam_psi, box ? &b : nullptr, a);
    if (rc) return rc;
    rc = adjoint_launch(p, a, st);
    if (rc) return rc;
    if (n >= 1 && nprobe > 0) {
      probe_scatter_kernel<<<(nprobe + 255) / 256, 256, 0, st>>>(
          gl, probe_comp, (const long long*)probe_idx, w,
          probe_hist + (long long)(n - 1) * nprobe, scale, nprobe);
      YEE_HIP_CHECK(hipGetLastError());
    }
  }
  return YEE_OK;
}

}  // extern "C"
