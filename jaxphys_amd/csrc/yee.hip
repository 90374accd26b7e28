// jaxphys_amd: MI355X (gfx950) kernels of the 3D Yee FDTD step and the
// C ABI declared in include/yee.h.
//
// Reference path: src/neurosim/em/fdtd3d.py:280-379 (`step`), scanned by
// jax.lax.scan at fdtd3d.py:393.  One time step is two passes over HBM:
//   * the H half-step (fdtd3d.py:303-330; periodic :290-295) reads E and H,
//     writes H;
//   * the E half-step (fdtd3d.py:332-359; periodic :297-301) reads H, E and
//     the per-cell coefficient(s), writes E, and applies in the same pass the
//     sponge damping (fdtd3d.py:361-366) or the CPML psi recursion, then the
//     point sources (fdtd3d.py:368-369, insertion order).
//
// Mapping to CDNA4: a 256-thread block is one 64-wide x row segment (one
// wavefront) by 4 y rows; each thread owns a column (i, j) and marches a
// chunk of z planes, carrying the z neighbour of the previous plane in
// registers, so every field is streamed from HBM once per pass with
// coalesced x-fastest loads.  x and y neighbours are re-read from L1/L2.
// Arithmetic is float32 with contraction disabled (-ffp-contract=off) so that
// the operation order, and therefore the rounding, is that of the reference's
// float32 expressions.

#include "yee_internal.h"

#include <mutex>

namespace yee {

thread_local std::string g_err;

// One psi recursion + curl term:  psi = b psi + c d;  return d/kappa + psi.
__device__ __forceinline__ float cpml_term(float* __restrict__ ps, const float* __restrict__ tab,
                                           int n, int a, int kind, float d) {
  const float b = tab[(3 * kind + 0) * n + a];
  const float c = tab[(3 * kind + 1) * n + a];
  const float kinv = tab[(3 * kind + 2) * n + a];
  const float v = b * (*ps) + c * d;
  *ps = v;
  return d * kinv + v;
}

// ---------------------------------------------------------------------------
// H half-step: H -= ch * curl(E) (forward differences).
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(kThreads) void yee_h_kernel(GridArgs g, Tiling tl, int kz,
                                                        const float* __restrict__ ex,
                                                        const float* __restrict__ ey,
                                                        const float* __restrict__ ez,
                                                        float* __restrict__ hx,
                                                        float* __restrict__ hy,
                                                        float* __restrict__ hz) {
  int bx, by, bz;
  if (!tile_of(tl, bx, by, bz)) return;
  const int i = bx * kBX + (threadIdx.x % kBX);
  const int j = by * kBY + (threadIdx.x / kBX);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kz;
  const int k1 = min(k0 + kz, nz);
  const unsigned P = (unsigned)nx * ny;
  constexpr bool per = MODE == kPeriodic;
  const bool xin = per || i < nx - 1;
  const bool yin = per || j < ny - 1;
  const unsigned di = (per && i == nx - 1) ? -(unsigned)(nx - 1) : 1;
  const unsigned dj = (per && j == ny - 1) ? -(unsigned)(ny - 1) * nx : nx;
  const float ch = g.ch;
  unsigned idx = (unsigned)k0 * P + (unsigned)j * nx + i;
  float ex0 = ex[idx], ey0 = ey[idx];
  const int p = g.p;
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  const int sx = xs ? slab_index(i, nx, p) : 0;
  const int sy = ys ? slab_index(j, ny, p) : 0;
  // branch-free neighbour offsets: an out-of-range neighbour is read at the
  // cell itself and the update is then not selected
  const unsigned dic = xin ? di : 0u;
  const unsigned djc = yin ? dj : 0u;
  // software pipeline: the loads of plane k+1 are issued before the
  // arithmetic and stores of plane k
  auto dk_of = [&](int k) -> unsigned {
    return (per && k == nz - 1) ? -(unsigned)(nz - 1) * P : ((per || k < nz - 1) ? P : 0u);
  };
  float c_ez0, c_ex1, c_ey1, c_ezj, c_exj, c_ezi, c_eyi, c_hx0, c_hy0, c_hz0;
  {
    const unsigned dk = dk_of(k0);
    c_ez0 = ez[idx];
    c_ex1 = ex[idx + dk];
    c_ey1 = ey[idx + dk];
    c_ezj = ez[idx + djc];
    c_exj = ex[idx + djc];
    c_ezi = ez[idx + dic];
    c_eyi = ey[idx + dic];
    c_hx0 = hx[idx];
    c_hy0 = hy[idx];
    c_hz0 = hz[idx];
  }
  for (int k = k0; k < k1; ++k, idx += P) {
    float n_ez0 = 0.f, n_ex1 = 0.f, n_ey1 = 0.f, n_ezj = 0.f, n_exj = 0.f, n_ezi = 0.f,
          n_eyi = 0.f, n_hx0 = 0.f, n_hy0 = 0.f, n_hz0 = 0.f;
    if (k + 1 < k1) {
      const unsigned q = idx + P;
      const unsigned dk = dk_of(k + 1);
      n_ez0 = ez[q];
      n_ex1 = ex[q + dk];
      n_ey1 = ey[q + dk];
      n_ezj = ez[q + djc];
      n_exj = ex[q + djc];
      n_ezi = ez[q + dic];
      n_eyi = ey[q + dic];
      n_hx0 = hx[q];
      n_hy0 = hy[q];
      n_hz0 = hz[q];
    }
    const bool zin = per || k < nz - 1;
    const bool zs = MODE == kCpml && (k < p || k >= nz - p);
    const int sz = zs ? slab_index(k, nz, p) : 0;
    float ax = c_ezj - c_ez0, bx = c_ey1 - ey0;  // fdtd3d.py:305-312
    float ay = c_ex1 - ex0, by = c_ezi - c_ez0;  // fdtd3d.py:314-321
    float az = c_eyi - ey0, bz = c_exj - ex0;    // fdtd3d.py:323-330
    if (MODE == kCpml) {
      if (yin && zin) {
        if (ys) ax = cpml_term(g.psi[kHxy] + ((unsigned)(k * 2 * p + sy) * nx + i), g.cpy, ny, j, 1, ax);
        if (zs) bx = cpml_term(g.psi[kHxz] + ((unsigned)(sz * ny + j) * nx + i), g.cpz, nz, k, 1, bx);
      }
      if (xin && zin) {
        if (zs) ay = cpml_term(g.psi[kHyz] + ((unsigned)(sz * ny + j) * nx + i), g.cpz, nz, k, 1, ay);
        if (xs) by = cpml_term(g.psi[kHyx] + ((unsigned)k * ny + j) * 2 * p + sx, g.cpx, nx, i, 1, by);
      }
      if (xin && yin) {
        if (xs) az = cpml_term(g.psi[kHzx] + ((unsigned)k * ny + j) * 2 * p + sx, g.cpx, nx, i, 1, az);
        if (ys) bz = cpml_term(g.psi[kHzy] + ((unsigned)(k * 2 * p + sy) * nx + i), g.cpy, ny, j, 1, bz);
      }
    }
    hx[idx] = (yin && zin) ? c_hx0 - ch * (ax - bx) : c_hx0;
    hy[idx] = (xin && zin) ? c_hy0 - ch * (ay - by) : c_hy0;
    hz[idx] = (xin && yin) ? c_hz0 - ch * (az - bz) : c_hz0;
    ex0 = c_ex1;
    ey0 = c_ey1;
    c_ez0 = n_ez0; c_ex1 = n_ex1; c_ey1 = n_ey1; c_ezj = n_ezj; c_exj = n_exj;
    c_ezi = n_ezi; c_eyi = n_eyi; c_hx0 = n_hx0; c_hy0 = n_hy0; c_hz0 = n_hz0;
  }
}

// ---------------------------------------------------------------------------
// E half-step: E = [Ca] E + Cb * curl(H) (backward differences), then sponge
// damping or nothing, then the fused point sources.
// ---------------------------------------------------------------------------
template <int MODE, bool LOSSY, bool REC = false>
__global__ __launch_bounds__(kThreads) void yee_e_kernel(GridArgs g, SrcArgs s, Tiling tl, int kz,
                                                        float* __restrict__ ex,
                                                        float* __restrict__ ey,
                                                        float* __restrict__ ez,
                                                        const float* __restrict__ hx,
                                                        const float* __restrict__ hy,
                                                        const float* __restrict__ hz,
                                                        RecArgs r) {
  int bx, by, bz;
  if (!tile_of(tl, bx, by, bz)) return;  // (the whole workgroup)
  // material palette in LDS
  __shared__ float spal[512];
  if (g.mat) {
    for (int q = threadIdx.x; q < 512; q += kThreads) spal[q] = g.pal[q];
    __syncthreads();
  }
  const int i = bx * kBX + (threadIdx.x % kBX);
  const int j = by * kBY + (threadIdx.x / kBX);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i >= nx || j >= ny) return;
  const int k0 = bz * kz;
  const int k1 = min(k0 + kz, nz);
  const unsigned P = (unsigned)nx * ny;
  constexpr bool per = MODE == kPeriodic;
  const bool xin = per || i > 0;
  const bool yin = per || j > 0;
  const unsigned di = (per && i == 0) ? (unsigned)(nx - 1) : -1;
  const unsigned dj = (per && j == 0) ? (unsigned)(ny - 1) * nx : -(unsigned)nx;
  unsigned idx = (unsigned)k0 * P + (unsigned)j * nx + i;
  float hxm = 0.f, hym = 0.f;  // H at plane k-1
  if (k0 > 0) {
    hxm = hx[idx - P];
    hym = hy[idx - P];
  } else if (per) {
    hxm = hx[idx + (unsigned)(nz - 1) * P];
    hym = hy[idx + (unsigned)(nz - 1) * P];
  }
  const int p = g.p;
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  const int sx = xs ? slab_index(i, nx, p) : 0;
  const int sy = ys ? slab_index(j, ny, p) : 0;
  float sxy = 0.f;
  if (MODE == kSponge) sxy = g.ssx[i] + g.ssy[j];
  const unsigned dic = xin ? di : 0u;
  const unsigned djc = yin ? dj : 0u;
  // recording: this column's offsets in the box and in the probe map
  bool rb = false, rp = false;
  unsigned rl = 0, pl = 0;
  if (REC) {
    const unsigned li = (unsigned)(i - r.bx0), lj = (unsigned)(j - r.by0);
    rb = r.box_out && li < (unsigned)r.bnx && lj < (unsigned)r.bny;
    rl = lj * r.bnx + li;
    const unsigned pi = (unsigned)(i - r.px0), pj = (unsigned)(j - r.py0);
    rp = r.probe_out && pi < (unsigned)r.pnx && pj < (unsigned)r.pny;
    pl = pj * r.pnx + pi;
  }
  float c_hx0, c_hy0, c_hz0, c_hzj, c_hxj, c_hzi, c_hyi, c_cb, c_ca = 1.f, c_ex0, c_ey0, c_ez0;
  {
    c_hx0 = hx[idx];
    c_hy0 = hy[idx];
    c_hz0 = hz[idx];
    c_hzj = hz[idx + djc];
    c_hxj = hx[idx + djc];
    c_hzi = hz[idx + dic];
    c_hyi = hy[idx + dic];
    if (g.mat) {
      c_cb = __uint_as_float(g.mat[idx]);  // material id, looked up at use
    } else {
      c_cb = g.cb[idx];
      c_ca = LOSSY ? g.ca[idx] : 1.f;
    }
    c_ex0 = ex[idx];
    c_ey0 = ey[idx];
    c_ez0 = ez[idx];
  }
  for (int k = k0; k < k1; ++k, idx += P) {
    float n_hx0 = 0.f, n_hy0 = 0.f, n_hz0 = 0.f, n_hzj = 0.f, n_hxj = 0.f, n_hzi = 0.f,
          n_hyi = 0.f, n_cb = 0.f, n_ca = 1.f, n_ex0 = 0.f, n_ey0 = 0.f, n_ez0 = 0.f;
    if (k + 1 < k1) {
      const unsigned q = idx + P;
      n_hx0 = hx[q];
      n_hy0 = hy[q];
      n_hz0 = hz[q];
      n_hzj = hz[q + djc];
      n_hxj = hx[q + djc];
      n_hzi = hz[q + dic];
      n_hyi = hy[q + dic];
      if (g.mat) {
        n_cb = __uint_as_float(g.mat[q]);
      } else {
        n_cb = g.cb[q];
        n_ca = LOSSY ? g.ca[q] : 1.f;
      }
      n_ex0 = ex[q];
      n_ey0 = ey[q];
      n_ez0 = ez[q];
    }
    const bool zin = per || k > 0;
    const float hx0 = c_hx0, hy0 = c_hy0, hz0 = c_hz0;
    float cb = c_cb, ca = c_ca;
    if (g.mat) {
      const unsigned m = __float_as_uint(c_cb);
      cb = spal[m];
      ca = LOSSY ? spal[256 + m] : 1.f;
    }
    const float ex0 = c_ex0, ey0 = c_ey0, ez0 = c_ez0;
    const bool zs = MODE == kCpml && (k < p || k >= nz - p);
    const int sz = zs ? slab_index(k, nz, p) : 0;
    float ax = hz0 - c_hzj, bx = hy0 - hym;  // fdtd3d.py:334-341
    float ay = hx0 - hxm, by = hz0 - c_hzi;  // fdtd3d.py:343-350
    float az = hy0 - c_hyi, bz = hx0 - c_hxj;  // fdtd3d.py:352-359
    if (MODE == kCpml) {
      if (yin && zin) {
        if (ys) ax = cpml_term(g.psi[kExy] + ((unsigned)(k * 2 * p + sy) * nx + i), g.cpy, ny, j, 0, ax);
        if (zs) bx = cpml_term(g.psi[kExz] + ((unsigned)(sz * ny + j) * nx + i), g.cpz, nz, k, 0, bx);
      }
      if (xin && zin) {
        if (zs) ay = cpml_term(g.psi[kEyz] + ((unsigned)(sz * ny + j) * nx + i), g.cpz, nz, k, 0, ay);
        if (xs) by = cpml_term(g.psi[kEyx] + ((unsigned)k * ny + j) * 2 * p + sx, g.cpx, nx, i, 0, by);
      }
      if (xin && yin) {
        if (xs) az = cpml_term(g.psi[kEzx] + ((unsigned)k * ny + j) * 2 * p + sx, g.cpx, nx, i, 0, az);
        if (ys) bz = cpml_term(g.psi[kEzy] + ((unsigned)(k * 2 * p + sy) * nx + i), g.cpy, ny, j, 0, bz);
      }
    }
    c_hx0 = n_hx0; c_hy0 = n_hy0; c_hz0 = n_hz0; c_hzj = n_hzj; c_hxj = n_hxj; c_hzi = n_hzi;
    c_hyi = n_hyi; c_cb = n_cb; c_ca = n_ca; c_ex0 = n_ex0; c_ey0 = n_ey0; c_ez0 = n_ez0;
    float vx = ex0, vy = ey0, vz = ez0;
    if (yin && zin) vx = LOSSY ? ca * ex0 + cb * (ax - bx) : ex0 + cb * (ax - bx);
    if (xin && zin) vy = LOSSY ? ca * ey0 + cb * (ay - by) : ey0 + cb * (ay - by);
    if (xin && yin) vz = LOSSY ? ca * ez0 + cb * (az - bz) : ez0 + cb * (az - bz);
    if (MODE == kSponge) {  // fdtd3d.py:361-366, sigma_avg of :221
      const float sig = (sxy + g.ssz[k]) / 3.0f;
      if (sig != 0.f) {  // exp(-0) == 1 exactly: the multiply is a no-op
        const float d = expf(((-sig) * g.sdt) / g.seps0);
        vx *= d;
        vy *= d;
        vz *= d;
      }
    }
#pragma unroll
    for (int q = 0; q < YEE_FUSED_SOURCES; ++q) {  // fdtd3d.py:236-244
      if (q < s.nsrc && s.idx[q] == idx) {
        if (s.comp[q] == 0) vx += s.val[q];
        else if (s.comp[q] == 1) vy += s.val[q];
        else vz += s.val[q];
      }
    }
    ex[idx] = vx;
    ey[idx] = vy;
    ez[idx] = vz;
    if (REC) {
      const unsigned lk = (unsigned)(k - r.bz0);
      if (rb && lk < (unsigned)r.bnz) {
        // d = E^{n+1} - Ca E^n, the same float operations as the adjoint's
        // accumulation applies to the two states
        const unsigned l = lk * r.bny * r.bnx + rl;
        r.box_out[l] = vx - (LOSSY ? ca * ex0 : ex0);
        r.box_out[r.bc + l] = vy - (LOSSY ? ca * ey0 : ey0);
        r.box_out[2u * r.bc + l] = vz - (LOSSY ? ca * ez0 : ez0);
        if (LOSSY) {
          r.box_out[3u * r.bc + l] = ex0;
          r.box_out[4u * r.bc + l] = ey0;
          r.box_out[5u * r.bc + l] = ez0;
        }
      }
      const unsigned pk = (unsigned)(k - r.pz0);
      if (rp && pk < (unsigned)r.pnz) {
        const unsigned pc = (unsigned)r.pnx * r.pny * r.pnz;
        const unsigned l = pk * r.pny * r.pnx + pl;
        const int q0 = r.pmap[l], q1 = r.pmap[pc + l], q2 = r.pmap[2 * pc + l];
        if (q0 >= 0) r.probe_out[q0] = vx;
        if (q1 >= 0) r.probe_out[q1] = vy;
        if (q2 >= 0) r.probe_out[q2] = vz;
      }
    }
    hxm = hx0;
    hym = hy0;
  }
}

// Sources beyond the fused ones: one thread, insertion order.
__global__ void yee_inject_kernel(float* ex, float* ey, float* ez, const int* comp,
                                  const long long* idx, const float* table, int nsrc,
                                  int first, int n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (int q = first; q < nsrc; ++q) {
    const float v = table[(long long)n * nsrc + q];
    float* f = comp[q] == 0 ? ex : (comp[q] == 1 ? ey : ez);
    f[idx[q]] += v;
  }
}

// ---------------------------------------------------------------------------
// float64 energy sums (deterministic two-stage reduction)
// ---------------------------------------------------------------------------
constexpr int kEnergyBlocks = 1024;

__device__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x / 64); ++w) r += sh[w];
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(256) void energy_partial_kernel(GridArgs g, const float* eps,
                                                            long long ncell, double* part) {
  __shared__ double sh[4];
  double we = 0.0, wh = 0.0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const double e = (double)g.f[0][c] * g.f[0][c] + (double)g.f[1][c] * g.f[1][c] +
                     (double)g.f[2][c] * g.f[2][c];
    we += (eps ? (double)eps[c] : 1.0) * e;
    wh += (double)g.f[3][c] * g.f[3][c] + (double)g.f[4][c] * g.f[4][c] +
          (double)g.f[5][c] * g.f[5][c];
  }
  we = block_sum(we, sh);
  wh = block_sum(wh, sh);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = we;
    part[2 * blockIdx.x + 1] = wh;
  }
}

__global__ __launch_bounds__(256) void energy_final_kernel(const double* part, int nb, double* out) {
  __shared__ double sh[4];
  double we = 0.0, wh = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    we += part[2 * b];
    wh += part[2 * b + 1];
  }
  we = block_sum(we, sh);
  wh = block_sum(wh, sh);
  if (threadIdx.x == 0) {
    out[0] = we;
    out[1] = wh;
  }
}

// Poynting flux through the plane normal to `axis` halfway between E planes
// p and p+1 (where the two tangential H components live): per cell of the
// plane, (E_a x H)_n with each tangential E averaged over its two planes.
//   axis x: S = Ey Hz - Ez Hy,  axis y: S = Ez Hx - Ex Hz,  axis z: S = Ex Hy - Ey Hx
__global__ __launch_bounds__(256) void flux_partial_kernel(GridArgs g, int axis, int p,
                                                          double* part) {
  __shared__ double sh[4];
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  // plane cells (u, v): axis x -> (j, k), y -> (i, k), z -> (i, j)
  const int nu = axis == 0 ? ny : nx;
  const int nv = axis == 2 ? ny : nz;
  const long long ncell = (long long)nu * nv;
  // first / second tangential E and the H each pairs with (right-handed)
  const int ea = axis == 0 ? 1 : (axis == 1 ? 2 : 0);
  const int eb = axis == 0 ? 2 : (axis == 1 ? 0 : 1);
  const int ha = axis == 0 ? 5 : (axis == 1 ? 3 : 4);
  const int hb = axis == 0 ? 4 : (axis == 1 ? 5 : 3);
  const long long step = axis == 0 ? 1 : (axis == 1 ? (long long)nx : (long long)nx * ny);
  double acc = 0.0;
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    const int u = (int)(c % nu), v = (int)(c / nu);
    const int i = axis == 0 ? p : u;
    const int j = axis == 0 ? u : (axis == 1 ? p : v);
    const int k = axis == 2 ? p : v;
    const long long o = ((long long)k * ny + j) * nx + i;
    const double ua = 0.5 * ((double)g.f[ea][o] + (double)g.f[ea][o + step]);
    const double ub = 0.5 * ((double)g.f[eb][o] + (double)g.f[eb][o + step]);
    acc += ua * (double)g.f[ha][o] - ub * (double)g.f[hb][o];
  }
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(256) void flux_final_kernel(const double* part, int nb, double* out) {
  __shared__ double sh[4];
  double acc = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) acc += part[b];
  acc = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = acc;
}

__global__ __launch_bounds__(256) void copy_probe_kernel(const float4* __restrict__ src,
                                                         float4* __restrict__ dst, long long n4) {
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n4;
       c += (long long)gridDim.x * blockDim.x)
    dst[c] = src[c];
}

__global__ void copy_tail_kernel(const float* src, float* dst, long long from, long long n) {
  const long long c = from + (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n) dst[c] = src[c];
}

void free_sources(YeePlan* p) {
  if (p->d_comp) (void)hipFree(p->d_comp);
  if (p->d_idx) (void)hipFree(p->d_idx);
  if (p->d_table) (void)hipFree(p->d_table);
  p->d_comp = nullptr;
  p->d_idx = nullptr;
  p->d_table = nullptr;
}

int validate(const YeeGridDesc* d) {
  if (!d) return fail(YEE_EINVAL, "null grid descriptor");
  if (d->nx < 2 || d->ny < 2 || d->nz < 2)
    return fail(YEE_EINVAL, "grid must be at least 2 cells per axis");
  if (d->boundary < YEE_BOUNDARY_WALLS || d->boundary > YEE_BOUNDARY_CPML)
    return fail(YEE_EINVAL, "unknown boundary code");
  if (!d->cb) return fail(YEE_EINVAL, "cb (ce/eps_r) array is required");
  if ((d->mat == nullptr) != (d->palette == nullptr))
    return fail(YEE_EINVAL, "mat and palette go together");
  if (d->boundary == YEE_BOUNDARY_SPONGE && !d->sponge_sigma)
    return fail(YEE_EINVAL, "sponge boundary needs sponge_sigma");
  if (d->boundary == YEE_BOUNDARY_CPML) {
    if (d->npml < 1) return fail(YEE_EINVAL, "cpml needs npml >= 1");
    if (2 * d->npml > d->nx || 2 * d->npml > d->ny || 2 * d->npml > d->nz)
      return fail(YEE_EINVAL, "cpml slabs overlap: need 2*npml <= n on every axis");
    if (!d->cpml) return fail(YEE_EINVAL, "cpml needs coefficient tables");
  }
  return YEE_OK;
}

#define YEE_FIELDS(g) (g).f[0], (g).f[1], (g).f[2], (g).f[3], (g).f[4], (g).f[5]

// two-pass kernels: XCD-aware tile order unless plain order is asked for
static Tiling plan_tiling(const YeePlan* p) {
  return make_tiling(p->d.nx, p->d.ny, p->d.nz, p->kz, p->xcd);
}

int launch_h(const YeePlan* p, const GridArgs& g, hipStream_t st) {
  const Tiling tl = plan_tiling(p);
  const unsigned nb = tiling_blocks(tl);
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC:
      yee_h_kernel<kPeriodic><<<nb, kThreads, 0, st>>>(g, tl, p->kz, YEE_FIELDS(g));
      break;
    case YEE_BOUNDARY_CPML:
      yee_h_kernel<kCpml><<<nb, kThreads, 0, st>>>(g, tl, p->kz, YEE_FIELDS(g));
      break;
    default:  // walls and sponge share the H update (sponge damps E only)
      yee_h_kernel<kWalls><<<nb, kThreads, 0, st>>>(g, tl, p->kz, YEE_FIELDS(g));
      break;
  }
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

template <bool LOSSY, bool REC>
void launch_e_mode(const YeePlan* p, const GridArgs& g, const SrcArgs& s, const RecArgs& r,
                   hipStream_t st) {
  const Tiling tl = plan_tiling(p);
  const unsigned nb = tiling_blocks(tl);
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC:
      yee_e_kernel<kPeriodic, LOSSY, REC><<<nb, kThreads, 0, st>>>(g, s, tl, p->kz,
                                                                   YEE_FIELDS(g), r);
      break;
    case YEE_BOUNDARY_CPML:
      yee_e_kernel<kCpml, LOSSY, REC><<<nb, kThreads, 0, st>>>(g, s, tl, p->kz, YEE_FIELDS(g),
                                                               r);
      break;
    case YEE_BOUNDARY_SPONGE:
      yee_e_kernel<kSponge, LOSSY, REC><<<nb, kThreads, 0, st>>>(g, s, tl, p->kz,
                                                                 YEE_FIELDS(g), r);
      break;
    default:
      yee_e_kernel<kWalls, LOSSY, REC><<<nb, kThreads, 0, st>>>(g, s, tl, p->kz, YEE_FIELDS(g),
                                                                r);
      break;
  }
}

int launch_e(const YeePlan* p, const GridArgs& g, int n, hipStream_t st, const RecArgs* rec) {
  SrcArgs s;
  int rc = make_src_args(p, n, s);
  if (rc) return rc;
  if (rec && p->nsrc > YEE_FUSED_SOURCES)
    return fail(YEE_EINVAL, "recording inside the E kernel needs every source fused");
  RecArgs r;
  std::memset(&r, 0, sizeof(r));
  if (rec) r = *rec;
  if (p->d.ca) {
    if (rec) launch_e_mode<true, true>(p, g, s, r, st);
    else launch_e_mode<true, false>(p, g, s, r, st);
  } else {
    if (rec) launch_e_mode<false, true>(p, g, s, r, st);
    else launch_e_mode<false, false>(p, g, s, r, st);
  }
  YEE_HIP_CHECK(hipGetLastError());
  return launch_inject(p, g.f[0], g.f[1], g.f[2], n, YEE_FUSED_SOURCES, st);
}

int launch_inject(const YeePlan* p, float* ex, float* ey, float* ez, int n, int first,
                  hipStream_t st) {
  if (p->nsrc > first) {
    yee_inject_kernel<<<1, 64, 0, st>>>(ex, ey, ez, p->d_comp, p->d_idx, p->d_table, p->nsrc,
                                        first, n);
    YEE_HIP_CHECK(hipGetLastError());
  }
  return YEE_OK;
}

// ---------------------------------------------------------------------------
// Workspace of yee_energy_fp64_noalloc: the float64 partial sums of the
// energy reduction in a file-scope device array instead of a caller buffer.
// Every GPU the library runs on holds its own copy (16 KiB) until the
// process ends.  In host code `scratch` names a host-side stand-in, never
// the device memory: the device address is looked up once per device with
// hipGetSymbolAddress and only that address is passed to kernels.
// ---------------------------------------------------------------------------
constexpr int kScratchDoubles = 2 * kEnergyBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on the
// first call per device, with that device current)
static int scratch_on_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(YEE_EHIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return fail(YEE_EUNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(YEE_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = (double*)ptr;
  }
  *out = g_scratch_ptr[dev];
  return YEE_OK;
}

}  // namespace yee

using namespace yee;

extern "C" {

const char* yee_last_error(void) { return g_err.c_str(); }

int64_t yee_psi_floats(const YeeGridDesc* d) {
  if (!d || d->boundary != YEE_BOUNDARY_CPML) return 0;
  return 4 * (psi_slab_floats(*d, 0) + psi_slab_floats(*d, 1) + psi_slab_floats(*d, 2));
}

int yee_plan_create(const YeeGridDesc* desc, YeePlan** out) {
  if (!out) return fail(YEE_EINVAL, "null output pointer");
  *out = nullptr;
  int rc = validate(desc);
  if (rc) return rc;
  // kernels use 32-bit cell offsets
  if ((long long)desc->nx * desc->ny * desc->nz >= (1LL << 32) ||
      yee_psi_floats(desc) >= (1LL << 32))
    return fail(YEE_EUNSUPPORTED, "grids of 2^32 cells or more are not supported");
  YeePlan* p = new (std::nothrow) YeePlan();
  if (!p) return fail(YEE_EINVAL, "out of host memory");
  p->d = *desc;
  // z chunking of the two-pass kernels: short column marches (many tiles);
  // the carried plane is re-read once per chunk from L2.  4 planes measured
  // fastest on the 512^3 grid (131072 tiles: 2.305 against 2.347 ms per
  // step with 2), 2 planes where 4 would leave fewer than 8192 tiles (160^3:
  // 72.9 against 75.3 us per step).
  int kz = 0;
  if (const char* e = std::getenv("JAXPHYS_AMD_KZ")) kz = std::atoi(e);
  if (kz <= 0) {
    const long long tiles4 = (long long)((desc->nx + 63) / 64) * ((desc->ny + 3) / 4) *
                             ((desc->nz + 3) / 4);
    kz = tiles4 >= 8192 ? 4 : 2;
  }
  if (kz > desc->nz) kz = desc->nz;
  p->kz = kz;
  if (const char* e = std::getenv("JAXPHYS_AMD_ADJ_KZ")) {
    const int v = std::atoi(e);
    if (v > 0) p->adj_kz = v;
  }
  if (p->adj_kz > desc->nz) p->adj_kz = desc->nz;
  if (const char* e = std::getenv("JAXPHYS_AMD_FUSED_KZ")) p->fused_kz = std::atoi(e);
  if (const char* e = std::getenv("JAXPHYS_AMD_ADJ_PIPE")) p->adj_pipe = std::atoi(e) != 0;
  if (const char* e = std::getenv("JAXPHYS_AMD_FUSED_VARIANT")) {
    const int v = std::atoi(e);
    if (v >= 0 && v <= 4) p->fused_variant = v;
  }
  if (const char* e = std::getenv("JAXPHYS_AMD_XCD")) {
    const int v = std::atoi(e);
    if (v >= -1 && v <= 10) p->xcd = v;
  }
  if (const char* e = std::getenv("JAXPHYS_AMD_XCD_SHELL")) {
    const int v = std::atoi(e);
    if (v >= -1 && v <= 10) p->xcd_shell = v;
  }
  if (const char* e = std::getenv("JAXPHYS_AMD_FUSED_ROWS")) {
    const int r = std::atoi(e);
    if (r == 4 || r == 8 || r == 16) p->fused_rows = r;
  }
  *out = p;
  return YEE_OK;
}

int yee_plan_tune(YeePlan* p, int32_t kz, int32_t fused_kz) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (kz > 0) p->kz = kz > p->d.nz ? p->d.nz : kz;
  if (fused_kz >= 0) p->fused_kz = fused_kz;
  return YEE_OK;
}

int yee_plan_set_variant(YeePlan* p, int32_t variant) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (variant < 0 || variant > 4) return fail(YEE_EINVAL, "unknown fused variant");
  p->fused_variant = variant;
  return YEE_OK;
}

int yee_plan_set_tile_order(YeePlan* p, int32_t xcd_aware) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (xcd_aware < -1 || xcd_aware > 10) return fail(YEE_EINVAL, "tile order must be in -1 .. 10");
  p->xcd = xcd_aware;
  return YEE_OK;
}

int yee_plan_set_palette(YeePlan* p, const uint8_t* mat, const float* palette) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if ((mat == nullptr) != (palette == nullptr))
    return fail(YEE_EINVAL, "mat and palette go together");
  p->d.mat = mat;
  p->d.palette = palette;
  return YEE_OK;
}

int yee_plan_set_fused_rows(YeePlan* p, int32_t rows) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (rows != 4 && rows != 8 && rows != 16) return fail(YEE_EINVAL, "fused rows must be 4, 8 or 16");
  p->fused_rows = rows;
  return YEE_OK;
}

void yee_plan_destroy(YeePlan* p) {
  if (!p) return;
  free_sources(p);
  if (p->d_adjsets) (void)hipFree(p->d_adjsets);
  delete p;
}

int yee_set_sources(YeePlan* p, int32_t nsrc, const int32_t* comp, const int32_t* pos,
                    const float* table, int32_t nsteps) {
  if (!p) return fail(YEE_EINVAL, "null plan");
  if (nsrc < 0 || nsteps < 0) return fail(YEE_EINVAL, "negative source or step count");
  if (nsrc > 0 && (!comp || !pos || (nsteps > 0 && !table)))
    return fail(YEE_EINVAL, "null source arrays");
  // every source is checked before the plan changes: a rejected call leaves
  // the previous sources in place
  std::vector<int> new_comp(nsrc, 0);
  std::vector<long long> new_idx(nsrc, 0);
  for (int q = 0; q < nsrc; ++q) {
    const int i = pos[3 * q], j = pos[3 * q + 1], k = pos[3 * q + 2];
    if (i < 0 || i >= p->d.nx || j < 0 || j >= p->d.ny || k < 0 || k >= p->d.nz)
      return fail(YEE_EINVAL, "source position out of grid bounds");
    if (comp[q] < YEE_EX || comp[q] > YEE_EZ)
      return fail(YEE_EINVAL, "source component must be EX, EY or EZ");
    new_comp[q] = comp[q];
    new_idx[q] = ((long long)k * p->d.ny + j) * p->d.nx + i;
  }
  std::vector<float> new_table;
  if (table) new_table.assign(table, table + (size_t)nsrc * nsteps);
  // from here on the old device copies are gone: until the new ones are in
  // place the plan has no sources, so a failure below leaves a plan that
  // runs without sources rather than one that points at freed memory
  free_sources(p);
  p->nsrc = 0;
  p->nsteps_table = 0;
  p->comp.clear();
  p->idx.clear();
  p->table.clear();
  if (nsrc > 0) {  // device copies: the fused step injects every source itself
    int* dc = nullptr;
    long long* di = nullptr;
    float* dt = nullptr;
    const size_t tab_floats = (size_t)nsrc * (nsteps > 0 ? nsteps : 1);
    hipError_t e = hipMalloc(&dc, sizeof(int) * nsrc);
    if (e == hipSuccess) e = hipMalloc(&di, sizeof(long long) * nsrc);
    if (e == hipSuccess) e = hipMalloc(&dt, sizeof(float) * tab_floats);
    if (e == hipSuccess)
      e = hipMemcpy(dc, new_comp.data(), sizeof(int) * nsrc, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(di, new_idx.data(), sizeof(long long) * nsrc, hipMemcpyHostToDevice);
    if (e == hipSuccess && nsteps > 0)
      e = hipMemcpy(dt, new_table.data(), sizeof(float) * (size_t)nsrc * nsteps,
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      if (dc) (void)hipFree(dc);
      if (di) (void)hipFree(di);
      if (dt) (void)hipFree(dt);
      return fail(YEE_EHIP, std::string("source table upload: ") + hipGetErrorString(e));
    }
    p->d_comp = dc;
    p->d_idx = di;
    p->d_table = dt;
  }
  p->comp.swap(new_comp);
  p->idx.swap(new_idx);
  p->table.swap(new_table);
  p->nsrc = nsrc;
  p->nsteps_table = nsteps;
  return YEE_OK;
}

int yee_step_h(YeePlan* p, float* const* fields, float* psi, void* stream) {
  int rc = check_fields(p, fields, psi);
  if (rc) return rc;
  return launch_h(p, make_args(p, fields, psi), (hipStream_t)stream);
}

int yee_step_e(YeePlan* p, float* const* fields, float* psi, int32_t n, void* stream) {
  int rc = check_fields(p, fields, psi);
  if (rc) return rc;
  return launch_e(p, make_args(p, fields, psi), n, (hipStream_t)stream);
}

int yee_run(YeePlan* p, float* const* fields, float* psi, int32_t n0, int32_t nsteps,
            void* stream) {
  int rc = check_fields(p, fields, psi);
  if (rc) return rc;
  if (nsteps < 0) return fail(YEE_EINVAL, "negative step count");
  const GridArgs g = make_args(p, fields, psi);
  hipStream_t st = (hipStream_t)stream;
  for (int n = n0; n < n0 + nsteps; ++n) {
    rc = launch_h(p, g, st);
    if (rc) return rc;
    rc = launch_e(p, g, n, st);
    if (rc) return rc;
  }
  return YEE_OK;
}

int yee_energy_fp64(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* eps_r,
                    double* work, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(YEE_EINVAL, "bad grid size");
  if (!fields || !work) return fail(YEE_EINVAL, "null pointer");
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  for (int c = 0; c < 6; ++c) {
    if (!fields[c]) return fail(YEE_EINVAL, "null field pointer");
    g.f[c] = fields[c];
  }
  const long long ncell = (long long)nx * ny * nz;
  hipStream_t st = (hipStream_t)stream;
  energy_partial_kernel<<<kEnergyBlocks, 256, 0, st>>>(g, eps_r, ncell, work + 2);
  YEE_HIP_CHECK(hipGetLastError());
  energy_final_kernel<<<1, 256, 0, st>>>(work + 2, kEnergyBlocks, work);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_energy_fp64_noalloc(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                            const float* eps_r, double* out, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(YEE_EINVAL, "bad grid size");
  if (!fields || !out) return fail(YEE_EINVAL, "null pointer");
  const long long ncell = (long long)nx * ny * nz;
  if (ncell > YEE_NOALLOC_MAX_CELLS)
    return fail(YEE_EINVAL, "yee_energy_fp64_noalloc: grid of " + std::to_string(ncell) +
                                " cells exceeds its maximum of " +
                                std::to_string((long long)YEE_NOALLOC_MAX_CELLS) + " cells");
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  for (int c = 0; c < 6; ++c) {
    if (!fields[c]) return fail(YEE_EINVAL, "null field pointer");
    g.f[c] = fields[c];
  }
  double* part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc) return rc;
  static_assert(2 * kEnergyBlocks <= kScratchDoubles, "energy partials exceed the scratch array");
  hipStream_t st = (hipStream_t)stream;
  // every partial slot is written by its block before the final kernel
  // reads it, so nothing left by an earlier call is read
  energy_partial_kernel<<<kEnergyBlocks, 256, 0, st>>>(g, eps_r, ncell, part);
  YEE_HIP_CHECK(hipGetLastError());
  energy_final_kernel<<<1, 256, 0, st>>>(part, kEnergyBlocks, out);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_poynting_flux_fp64(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           int32_t axis, int32_t plane, double* work, void* stream) {
  if (nx < 1 || ny < 1 || nz < 1) return fail(YEE_EINVAL, "bad grid size");
  if (!fields || !work) return fail(YEE_EINVAL, "null pointer");
  if (axis < 0 || axis > 2) return fail(YEE_EINVAL, "axis must be 0, 1 or 2");
  const int n_axis = axis == 0 ? nx : (axis == 1 ? ny : nz);
  if (plane < 0 || plane + 1 >= n_axis)
    return fail(YEE_EINVAL, "plane must satisfy 0 <= plane < n - 1 along the axis");
  GridArgs g;
  std::memset(&g, 0, sizeof(g));
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  for (int c = 0; c < 6; ++c) {
    if (!fields[c]) return fail(YEE_EINVAL, "null field pointer");
    g.f[c] = fields[c];
  }
  hipStream_t st = (hipStream_t)stream;
  flux_partial_kernel<<<kEnergyBlocks, 256, 0, st>>>(g, axis, plane, work + 1);
  YEE_HIP_CHECK(hipGetLastError());
  flux_final_kernel<<<1, 256, 0, st>>>(work + 1, kEnergyBlocks, work);
  YEE_HIP_CHECK(hipGetLastError());
  return YEE_OK;
}

int yee_copy_probe(const float* src, float* dst, int64_t n, void* stream) {
  if (!src || !dst || n < 0) return fail(YEE_EINVAL, "bad copy arguments");
  hipStream_t st = (hipStream_t)stream;
  const long long n4 = n / 4;
  if (((uintptr_t)src | (uintptr_t)dst) % 16 != 0) return fail(YEE_EINVAL, "copy needs 16-byte alignment");
  if (n4 > 0) {
    long long blocks = (n4 + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    copy_probe_kernel<<<(unsigned)blocks, 256, 0, st>>>((const float4*)src, (float4*)dst, n4);
    YEE_HIP_CHECK(hipGetLastError());
  }
  if (n > n4 * 4) {
    copy_tail_kernel<<<1, 64, 0, st>>>(src, dst, n4 * 4, n);
    YEE_HIP_CHECK(hipGetLastError());
  }
  return YEE_OK;
}

}  // extern "C"
