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
// H(x-1) for its row (the previous 64-wide tile's last column).  H(z-1) is
// the thread's own H from the previous plane (register carry); each block
// marches a chunk of z planes after recomputing the plane below the chunk.
// Per plane: compute own H -> LDS (double-buffered by plane parity) and to
// B; one barrier; owners compute E^{n+1} from LDS + registers -> B.
// Arithmetic and its order are those of the two-pass kernels, so the result
// is bit-for-bit the same.

#include "yee_internal.h"

namespace yee {

struct FusedArgs {
  GridArgs in;            // input fields / psi slots + all coefficients
  float* fo[6];           // output fields
  float* po[kNumPsi];     // output psi slots
};

template <int MODE>
__device__ __forceinline__ float cpml_in_out(const float* pin, float* pout, bool write,
                                             const float* tab, int n, int v, int kind, float d) {
  const float b = tab[(3 * kind + 0) * n + v];
  const float c = tab[(3 * kind + 1) * n + v];
  const float kinv = tab[(3 * kind + 2) * n + v];
  const float ps = b * (*pin) + c * d;
  if (write) *pout = ps;
  return d * kinv + ps;
}

// H^{n+1/2} at (i, j, k) (indices already wrapped/valid) from buffer A;
// writes psi_H' to buffer B when `own`.
template <int MODE>
__device__ __forceinline__ void h_cell(const FusedArgs& a, int i, int j, int k, bool own,
                                       float& ox, float& oy, float& oz) {
  const GridArgs& g = a.in;
  const int nx = g.nx, ny = g.ny, nz = g.nz, p = g.p;
  constexpr bool per = MODE == kPeriodic;
  const long long P = (long long)nx * ny;
  const long long c = ((long long)k * ny + j) * nx + i;
  const bool xin = per || i < nx - 1;
  const bool yin = per || j < ny - 1;
  const bool zin = per || k < nz - 1;
  const long long di = (per && i == nx - 1) ? -(long long)(nx - 1) : 1;
  const long long dj = (per && j == ny - 1) ? -(long long)(ny - 1) * nx : nx;
  const long long dk = (per && k == nz - 1) ? -(long long)(nz - 1) * P : P;
  const float* __restrict__ ex = g.f[0];
  const float* __restrict__ ey = g.f[1];
  const float* __restrict__ ez = g.f[2];
  const float ex0 = ex[c], ey0 = ey[c], ez0 = ez[c];
  float hx = g.f[3][c], hy = g.f[4][c], hz = g.f[5][c];
  float ex1 = 0.f, ey1 = 0.f, ezj = 0.f, exj = 0.f, ezi = 0.f, eyi = 0.f;
  if (zin) { ex1 = ex[c + dk]; ey1 = ey[c + dk]; }
  if (yin) { ezj = ez[c + dj]; exj = ex[c + dj]; }
  if (xin) { ezi = ez[c + di]; eyi = ey[c + di]; }
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  const bool zs = MODE == kCpml && (k < p || k >= nz - p);
  const long long oy_ = ys ? ((long long)(k * 2 * p + slab_index(j, ny, p)) * nx + i) : 0;
  const long long oz_ = zs ? ((long long)(slab_index(k, nz, p) * ny + j) * nx + i) : 0;
  const long long ox_ = xs ? (((long long)k * ny + j) * 2 * p + slab_index(i, nx, p)) : 0;
  const float ch = g.ch;
  if (yin && zin) {  // fdtd3d.py:305-312
    float u = ezj - ez0, v = ey1 - ey0;
    if (MODE == kCpml) {
      if (ys) u = cpml_in_out<MODE>(g.psi[kHxy] + oy_, a.po[kHxy] + oy_, own, g.cpy, ny, j, 1, u);
      if (zs) v = cpml_in_out<MODE>(g.psi[kHxz] + oz_, a.po[kHxz] + oz_, own, g.cpz, nz, k, 1, v);
    }
    hx = hx - ch * (u - v);
  }
  if (xin && zin) {  // fdtd3d.py:314-321
    float u = ex1 - ex0, v = ezi - ez0;
    if (MODE == kCpml) {
      if (zs) u = cpml_in_out<MODE>(g.psi[kHyz] + oz_, a.po[kHyz] + oz_, own, g.cpz, nz, k, 1, u);
      if (xs) v = cpml_in_out<MODE>(g.psi[kHyx] + ox_, a.po[kHyx] + ox_, own, g.cpx, nx, i, 1, v);
    }
    hy = hy - ch * (u - v);
  }
  if (xin && yin) {  // fdtd3d.py:323-330
    float u = eyi - ey0, v = exj - ex0;
    if (MODE == kCpml) {
      if (xs) u = cpml_in_out<MODE>(g.psi[kHzx] + ox_, a.po[kHzx] + ox_, own, g.cpx, nx, i, 1, u);
      if (ys) v = cpml_in_out<MODE>(g.psi[kHzy] + oy_, a.po[kHzy] + oy_, own, g.cpy, ny, j, 1, v);
    }
    hz = hz - ch * (u - v);
  }
  ox = hx;
  oy = hy;
  oz = hz;
}

template <int MODE, bool LOSSY, int BY>
__global__ __launch_bounds__(64 * (BY + 1)) void yee_fused_kernel(FusedArgs a, SrcArgs s, int kz) {
  __shared__ float shx[2][BY + 1][64];
  __shared__ float shy[2][BY + 1][64];
  __shared__ float shz[2][BY + 1][64];
  const GridArgs& g = a.in;
  const int nx = g.nx, ny = g.ny, nz = g.nz, p = g.p;
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
  const long long P = (long long)nx * ny;
  // H(z-1) of the first plane: recompute the plane below the chunk
  float hxm = 0.f, hym = 0.f;
  {
    int kp = k0 - 1;
    if (per && kp < 0) kp += nz;
    if (owner && kp >= 0) {
      float tz;
      h_cell<MODE>(a, i, j, kp, false, hxm, hym, tz);
    }
  }
  const bool xs = MODE == kCpml && (i < p || i >= nx - p);
  const bool ys = MODE == kCpml && (j < p || j >= ny - p);
  float sxy = 0.f;
  if (MODE == kSponge && owner) sxy = g.ssx[i] + g.ssy[j];
  for (int k = k0; k < k1; ++k) {
    const int buf = k & 1;
    float hx = 0.f, hy = 0.f, hz = 0.f;
    if (valid) h_cell<MODE>(a, i, j, k, owner, hx, hy, hz);
    shx[buf][w][lane] = hx;
    shy[buf][w][lane] = hy;
    shz[buf][w][lane] = hz;
    float hyxm = 0.f, hzxm = 0.f;
    if (owner && lane == 0 && imvalid) {
      float tx;
      h_cell<MODE>(a, im, j, k, false, tx, hyxm, hzxm);
    }
    const long long c = (long long)k * P + (long long)j * nx + i;
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
      const float cb = g.cb[c];
      const float ca = LOSSY ? g.ca[c] : 1.f;
      float vx = g.f[0][c], vy = g.f[1][c], vz = g.f[2][c];
      const bool zs = MODE == kCpml && (k < p || k >= nz - p);
      const long long oy_ = ys ? ((long long)(k * 2 * p + slab_index(j, ny, p)) * nx + i) : 0;
      const long long oz_ = zs ? ((long long)(slab_index(k, nz, p) * ny + j) * nx + i) : 0;
      const long long ox_ = xs ? (((long long)k * ny + j) * 2 * p + slab_index(i, nx, p)) : 0;
      if (yin && zin) {  // fdtd3d.py:334-341
        float u = hz - hzj, v = hy - hym;
        if (MODE == kCpml) {
          if (ys) u = cpml_in_out<MODE>(g.psi[kExy] + oy_, a.po[kExy] + oy_, true, g.cpy, ny, j, 0, u);
          if (zs) v = cpml_in_out<MODE>(g.psi[kExz] + oz_, a.po[kExz] + oz_, true, g.cpz, nz, k, 0, v);
        }
        vx = LOSSY ? ca * vx + cb * (u - v) : vx + cb * (u - v);
      }
      if (xin && zin) {  // fdtd3d.py:343-350
        float u = hx - hxm, v = hz - hzi;
        if (MODE == kCpml) {
          if (zs) u = cpml_in_out<MODE>(g.psi[kEyz] + oz_, a.po[kEyz] + oz_, true, g.cpz, nz, k, 0, u);
          if (xs) v = cpml_in_out<MODE>(g.psi[kEyx] + ox_, a.po[kEyx] + ox_, true, g.cpx, nx, i, 0, v);
        }
        vy = LOSSY ? ca * vy + cb * (u - v) : vy + cb * (u - v);
      }
      if (xin && yin) {  // fdtd3d.py:352-359
        float u = hy - hyi, v = hx - hxj;
        if (MODE == kCpml) {
          if (xs) u = cpml_in_out<MODE>(g.psi[kEzx] + ox_, a.po[kEzx] + ox_, true, g.cpx, nx, i, 0, u);
          if (ys) v = cpml_in_out<MODE>(g.psi[kEzy] + oy_, a.po[kEzy] + oy_, true, g.cpy, ny, j, 0, v);
        }
        vz = LOSSY ? ca * vz + cb * (u - v) : vz + cb * (u - v);
      }
      if (MODE == kSponge) {  // fdtd3d.py:361-366
        const float sig = (sxy + g.ssz[k]) / 3.0f;
        if (sig != 0.f) {
          const float d = expf(((-sig) * g.sdt) / g.seps0);
          vx *= d;
          vy *= d;
          vz *= d;
        }
      }
#pragma unroll
      for (int q = 0; q < YEE_FUSED_SOURCES; ++q) {  // fdtd3d.py:236-244
        if (q < s.nsrc && s.idx[q] == c) {
          if (s.comp[q] == 0) vx += s.val[q];
          else if (s.comp[q] == 1) vy += s.val[q];
          else vz += s.val[q];
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

constexpr int kFusedBY = 8;

int fused_kz(const YeePlan* p, int gx, int gy) {
  int kz = 0;
  if (const char* e = std::getenv("JAXPHYS_AMD_FUSED_KZ")) kz = std::atoi(e);
  if (kz <= 0) {
    const long long want = 2048;
    long long chunks = (want + (long long)gx * gy - 1) / ((long long)gx * gy);
    if (chunks < 1) chunks = 1;
    if (chunks > p->d.nz) chunks = p->d.nz;
    kz = (int)((p->d.nz + chunks - 1) / chunks);
  }
  return kz > p->d.nz ? p->d.nz : kz;
}

template <int MODE, bool LOSSY>
void launch_fused_t(const FusedArgs& a, const SrcArgs& s, dim3 grid, int kz, hipStream_t st) {
  yee_fused_kernel<MODE, LOSSY, kFusedBY><<<grid, 64 * (kFusedBY + 1), 0, st>>>(a, s, kz);
}

int launch_fused(const YeePlan* p, float* const* fin, float* const* fout, float* pin, float* pout,
                 int n, hipStream_t st) {
  SrcArgs s;
  int rc = make_src_args(p, n, s);
  if (rc) return rc;
  FusedArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = make_args(p, fin, pin);
  const GridArgs go = make_args(p, fout, pout);
  for (int c = 0; c < 6; ++c) a.fo[c] = fout[c];
  for (int q = 0; q < kNumPsi; ++q) a.po[q] = go.psi[q];
  const int gx = (p->d.nx + 63) / 64;
  const int gy = (p->d.ny + kFusedBY - 1) / kFusedBY;
  const int kz = fused_kz(p, gx, gy);
  const dim3 grid(gx, gy, (p->d.nz + kz - 1) / kz);
  const bool lossy = p->d.ca != nullptr;
  switch (p->d.boundary) {
    case YEE_BOUNDARY_PERIODIC:
      lossy ? launch_fused_t<kPeriodic, true>(a, s, grid, kz, st)
            : launch_fused_t<kPeriodic, false>(a, s, grid, kz, st);
      break;
    case YEE_BOUNDARY_CPML:
      lossy ? launch_fused_t<kCpml, true>(a, s, grid, kz, st)
            : launch_fused_t<kCpml, false>(a, s, grid, kz, st);
      break;
    case YEE_BOUNDARY_SPONGE:
      lossy ? launch_fused_t<kSponge, true>(a, s, grid, kz, st)
            : launch_fused_t<kSponge, false>(a, s, grid, kz, st);
      break;
    default:
      lossy ? launch_fused_t<kWalls, true>(a, s, grid, kz, st)
            : launch_fused_t<kWalls, false>(a, s, grid, kz, st);
      break;
  }
  YEE_HIP_CHECK(hipGetLastError());
  return launch_inject(p, go, n, st);
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
