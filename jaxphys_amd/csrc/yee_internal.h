// Internal definitions shared by the forward (yee.hip) and adjoint
// (yee_adjoint.hip) translation units of libyee_hip.so.
#pragma once

#include <hip/hip_runtime.h>

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
};

struct SrcArgs {
  int nsrc;  // sources handled inside the E kernel (<= YEE_FUSED_SOURCES)
  int comp[YEE_FUSED_SOURCES];
  long long idx[YEE_FUSED_SOURCES];
  float val[YEE_FUSED_SOURCES];
};

inline __device__ int slab_index(int a, int n, int p) {
  return a < p ? a : a - (n - 2 * p);
}

}  // namespace yee

// ---------------------------------------------------------------------------
// plan
// ---------------------------------------------------------------------------
struct YeePlan {
  YeeGridDesc d;
  int kz;            // z planes per block of the two-pass kernels
  int fused_kz = 0;  // z planes per block of the fused kernel (0: automatic)
  int fused_variant = 0;  // 0: LDS + halo row, 1: barrier-free recompute
  dim3 grid;
  // sources
  int nsrc = 0;
  int nsteps_table = 0;
  std::vector<int> comp;
  std::vector<long long> idx;
  std::vector<float> table;  // host copy, step-major
  int* d_comp = nullptr;     // device copies for the unfused tail
  long long* d_idx = nullptr;
  float* d_table = nullptr;
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
int launch_e(const YeePlan* p, const GridArgs& g, int n, hipStream_t st);

}  // namespace yee
