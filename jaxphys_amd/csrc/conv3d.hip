// jaxphys_amd: BF16 matrix-core (MFMA) 3x3x3 convolutions for the learned
// dispersion closure, gfx950.  C ABI in include/conv3d.h.  No reference
// counterpart (the reference has no neural network code).
//
// Forward (and input gradient, via spatially flipped / channel-swapped
// weights) is an implicit GEMM  D[co][p] = sum_k W[co][k] X[p + off(k)][c(k)]
// with M = output channels (16-row tiles), N = output positions (16
// consecutive x per tile), K = 27 taps x input channels (32 per MFMA).
// One v_mfma_f32_16x16x32_bf16 per (M tile, N tile, K step); float32
// accumulation.  The layer's weights are staged once per block in LDS; each
// wave owns one y row of 16*NT x positions and marches a chunk of z planes;
// input fragments are 16-byte channel-contiguous loads (channels last).
// Epilogues fuse bias, erf GELU (erf to 1.5e-7, see erf_as), the GELU
// derivative of the backward pass and the float32 output of the last layer.
//
// Weight gradient: dW[co][k] = sum_p dY[p][co] X[p + off(k)][c(k)], an MFMA
// GEMM with K = positions.  Each block reduces a set of 32-position row
// segments; dY and the 3x3 neighbourhood rows of X are staged transposed in
// LDS so that fragments are 8 consecutive positions; per-block partial sums
// go to a work buffer reduced by a second kernel in fixed order.
//
// MFMA register layout used (16x16x32 bf16): lane t holds A[t%16][8(t/16)+v]
// and B[8(t/16)+v][t%16] (v = 0..7) and D[4(t/16)+v][t%16] (v = 0..3).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/conv3d.h"

namespace conv {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define CONV_CHECK(expr)                                                      \
  do {                                                                        \
    hipError_t e_ = (expr);                                                   \
    if (e_ != hipSuccess) return fail(-2, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned short u16;
typedef unsigned short u16x4 __attribute__((ext_vector_type(4)));
typedef int i32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((unsigned)h << 16); }
// float -> bf16, round to nearest even (v_cvt_pk_bf16_f32 on gfx950)
__device__ __forceinline__ u16 f2bf(float f) {
  const __bf16 h = (__bf16)f;
  u16 u;
  __builtin_memcpy(&u, &h, 2);
  return u;
}

__device__ __forceinline__ bf16x8 load8(const u16* p) {
  uint4 v = *reinterpret_cast<const uint4*>(p);
  bf16x8 r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// erf by Abramowitz & Stegun, Handbook of Mathematical Functions (1964),
// formula 7.1.26: erf(x) = 1 - t (a1 + t (a2 + t (a3 + t (a4 + t a5))))
// exp(-x^2), t = 1 / (1 + p |x|), odd in x; |error| <= 1.5e-7, far below
// the bf16 rounding of the activations.  Branch-free: a reciprocal, five
// FMAs and one exp2 per value, where a library erff costs several times as
// many VALU cycles (the GELU epilogue competes with the MFMAs for issue).
__device__ __forceinline__ float erf_as(float x) {
  const float ax = fabsf(x);
  const float t = __builtin_amdgcn_rcpf(__builtin_fmaf(0.3275911f, ax, 1.0f));
  float q = __builtin_fmaf(t, 1.061405429f, -1.453152027f);
  q = __builtin_fmaf(t, q, 1.421413741f);
  q = __builtin_fmaf(t, q, -0.284496736f);
  q = __builtin_fmaf(t, q, 0.254829592f);
  const float y = __builtin_fmaf(-(q * t), __expf(-ax * ax), 1.0f);
  return __builtin_copysignf(y, x);
}
// GELU(z) = z Phi(z) = z (1 + erf(z / sqrt 2)) / 2 and its derivative
__device__ __forceinline__ float gelu(float z) {
  return 0.5f * z * (1.0f + erf_as(z * 0.70710678118654752f));
}
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erf_as(z * 0.70710678118654752f)) +
         z * 0.39894228040143268f * __expf(-0.5f * z * z);
}

// Cross-lane moves inside the 16-lane DPP rows, which are exactly the
// column groups of the 16x16 MFMA B operand (lane t holds column t % 16):
// shift_right gives lane r the value of lane r-1 and lane 0 the value of
// lane 15 of `prev`; shift_left gives lane r the value of lane r+1 and lane
// 15 the value of lane 0 of `next`.  Used to derive the x-1 / x+1 taps of
// a conv row from the x+0 fragments already in registers.
__device__ __forceinline__ bf16x8 shift_right(bf16x8 cur, bf16x8 prev) {
  int c[4], p[4], o[4];
  __builtin_memcpy(c, &cur, 16);
  __builtin_memcpy(p, &prev, 16);
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const int e = __builtin_amdgcn_update_dpp(0, p[d], 0x121 /* row_ror:1 */, 0xF, 0xF, false);
    o[d] = __builtin_amdgcn_update_dpp(e, c[d], 0x111 /* row_shr:1 */, 0xF, 0xF, false);
  }
  bf16x8 r;
  __builtin_memcpy(&r, o, 16);
  return r;
}
__device__ __forceinline__ bf16x8 shift_left(bf16x8 cur, bf16x8 next) {
  int c[4], n[4], o[4];
  __builtin_memcpy(c, &cur, 16);
  __builtin_memcpy(n, &next, 16);
#pragma unroll
  for (int d = 0; d < 4; ++d) {
    const int e = __builtin_amdgcn_update_dpp(0, n[d], 0x12F /* row_ror:15 */, 0xF, 0xF, false);
    o[d] = __builtin_amdgcn_update_dpp(e, c[d], 0x101 /* row_shl:1 */, 0xF, 0xF, false);
  }
  bf16x8 r;
  __builtin_memcpy(&r, o, 16);
  return r;
}

// One DPP move per dword: shift_in_right gives lane r the value of lane r-1
// of `cur` and lane 0 its own value of `edge`; shift_in_left gives lane r the
// value of lane r+1 of `cur` and lane 15 its own value of `edge` (DPP with
// bound_ctrl off keeps the old value where the source lane leaves the row).
__device__ __forceinline__ bf16x8 shift_in_right(bf16x8 cur, bf16x8 edge) {
  int c[4], e[4], o[4];
  __builtin_memcpy(c, &cur, 16);
  __builtin_memcpy(e, &edge, 16);
#pragma unroll
  for (int d = 0; d < 4; ++d)
    o[d] = __builtin_amdgcn_update_dpp(e[d], c[d], 0x111 /* row_shr:1 */, 0xF, 0xF, false);
  bf16x8 r;
  __builtin_memcpy(&r, o, 16);
  return r;
}
__device__ __forceinline__ bf16x8 shift_in_left(bf16x8 cur, bf16x8 edge) {
  int c[4], e[4], o[4];
  __builtin_memcpy(c, &cur, 16);
  __builtin_memcpy(e, &edge, 16);
#pragma unroll
  for (int d = 0; d < 4; ++d)
    o[d] = __builtin_amdgcn_update_dpp(e[d], c[d], 0x101 /* row_shl:1 */, 0xF, 0xF, false);
  bf16x8 r;
  __builtin_memcpy(&r, o, 16);
  return r;
}

struct ConvArgs {
  int nx, ny, nz, cin, cout, cout_pad, ycp, kpad, zchunk;
  const u16* x;
  const u16* w;
  const float* bias;
  u16* zout;
  const u16* aux;
  void* y;
};

// WAVES wavefronts per block share the staged weights (with 16 and up to
// 128 VGPRs per lane, one block fills a CU: 4 wavefronts per SIMD).  The
// input fragments of K step k+1 are loaded before the MFMAs of step k.
// SHIFT (input channels a multiple of 32, so a K step is one tap): the K
// loop walks the 9 (dz, dy) pairs; per pair and 32-channel block it loads
// the NT fragments of the x+0 tap plus one fragment each side, and makes the
// x-1 and x+1 fragments with shift_right / shift_left, so each input row is
// read once instead of three times.
// CIN > 0 fixes the input channel count at compile time (the divisions of
// the tap / channel split become shifts); 0 reads it from the arguments.
template <int EPI, int NT, int MT, int WAVES, bool SHIFT, int CIN = 0>
__global__ __launch_bounds__(64 * WAVES) void conv_fwd_kernel(ConvArgs a) {
  extern __shared__ u16 wsh[];  // [cout_pad][kpad + 8]
  const int ldw = a.kpad + 8;
  // stage weights
  {
    const int n16 = a.cout_pad * (a.kpad / 8);
    for (int q = threadIdx.x; q < n16; q += 64 * WAVES) {
      const int row = q / (a.kpad / 8), col = (q % (a.kpad / 8)) * 8;
      *reinterpret_cast<uint4*>(wsh + row * ldw + col) =
          *reinterpret_cast<const uint4*>(a.w + (long long)row * a.kpad + col);
    }
  }
  __syncthreads();
  const int t = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int g = t >> 4;       // k group
  const int r = t & 15;       // row / column within the tile
  const int x0 = blockIdx.x * 16 * NT;
  const int yy = blockIdx.y * WAVES + wv;
  const int z0 = blockIdx.z * a.zchunk;
  const int z1 = min(z0 + a.zchunk, a.nz);
  const int nx = a.nx, ny = a.ny, nz = a.nz, cin = CIN > 0 ? CIN : a.cin;
  const bool rowok = yy < ny;
  // 8-byte epilogue stores: channel pitch a multiple of 4, aligned outputs
  const bool vec4 = (a.ycp & 3) == 0 && ((uintptr_t)a.y & 7) == 0 &&
                    ((uintptr_t)a.zout & 7) == 0 && ((uintptr_t)a.aux & 7) == 0;
  // the NT input fragments of K step k0 at plane zz (zero outside the grid)
  auto load_b = [&](int zz, int k0, bf16x8* bf) {
    const int k = k0 + 8 * g;
    const int tap = k / cin;
    const int c0 = k - tap * cin;
    const int dz = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dx = tap % 3 - 1;
    const int sz = zz + dz, sy = yy + dy;
    const bool pl = tap < 27 && sz >= 0 && sz < nz && sy >= 0 && sy < ny;
    const u16* rowp = a.x + ((long long)(pl ? sz : 0) * ny + (pl ? sy : 0)) * nx * cin;
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int sx = x0 + n * 16 + r + dx;
      if (pl && sx >= 0 && sx < nx) bf[n] = load8(rowp + (long long)sx * cin + c0);
      else bf[n] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
    }
  };
  for (int zz = z0; zz < z1; ++zz) {
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    if (rowok && SHIFT) {
      const int ncb = cin >= 32 ? cin / 32 : 1;  // (SHIFT implies cin % 32 == 0)
      const int nsteps = 9 * ncb;
      // x+0 fragments of (pair p, channel block cb) at plane zz and the
      // fragments just left of the first / right of the last column tile
      auto load_c = [&](int st, bf16x8* c, bf16x8& el, bf16x8& er) {
        const int p = st / ncb, cb = st - p * ncb;
        const int sz = zz + p / 3 - 1, sy = yy + p % 3 - 1;
        const bool pl = sz >= 0 && sz < nz && sy >= 0 && sy < ny;
        const u16* rowp = a.x + ((long long)(pl ? sz : 0) * ny + (pl ? sy : 0)) * nx * cin +
                          cb * 32 + 8 * g;
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          const int sx = x0 + n * 16 + r;
          if (pl && sx < nx) c[n] = load8(rowp + (long long)sx * cin);
          else c[n] = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        }
        const int xl = x0 - 1, xr = x0 + 16 * NT;
        if (pl && xl >= 0) el = load8(rowp + (long long)xl * cin);
        else el = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
        if (pl && xr < nx) er = load8(rowp + (long long)xr * cin);
        else er = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
      };
      bf16x8 cur[NT], nxt[NT], el, er, nel, ner;
      load_c(0, cur, el, er);
      for (int st = 0; st < nsteps; ++st) {
        if (st + 1 < nsteps) load_c(st + 1, nxt, nel, ner);
        const int p = st / ncb, cb = st - p * ncb;
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          bf16x8 bfr[NT];
#pragma unroll
          for (int n = 0; n < NT; ++n) {
            if (dx == 0) bfr[n] = shift_right(cur[n], n == 0 ? el : cur[n - 1]);
            else if (dx == 1) bfr[n] = cur[n];
            else bfr[n] = shift_left(cur[n], n == NT - 1 ? er : cur[n + 1]);
          }
          const int k = (3 * p + dx) * cin + cb * 32 + 8 * g;
          bf16x8 af[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) af[m] = load8(wsh + (m * 16 + r) * ldw + k);
#pragma unroll
          for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int m = 0; m < MT; ++m)
              acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bfr[n], acc[m][n], 0, 0, 0);
        }
#pragma unroll
        for (int n = 0; n < NT; ++n) cur[n] = nxt[n];
        el = nel;
        er = ner;
      }
    } else if (rowok) {
      bf16x8 cur[NT], nxt[NT];
      load_b(zz, 0, cur);
      for (int k0 = 0; k0 < a.kpad; k0 += 32) {
        if (k0 + 32 < a.kpad) load_b(zz, k0 + 32, nxt);
        const int k = k0 + 8 * g;
        bf16x8 af[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) af[m] = load8(wsh + (m * 16 + r) * ldw + k);
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], cur[n], acc[m][n], 0, 0, 0);
#pragma unroll
        for (int n = 0; n < NT; ++n) cur[n] = nxt[n];
      }
    }
    if (!rowok) continue;
    // epilogue: lane holds D[co = m*16 + 4g + v][pos = x0 + n*16 + r]
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int px = x0 + n * 16 + r;
      if (px >= nx) continue;
      const long long pos = ((long long)zz * ny + yy) * nx + px;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        // bf16 outputs: the lane's 4 consecutive channels in one 8-byte
        // store (same values as the per-channel path below)
        if (EPI != CONV_EPI_LINEAR_F32 && vec4) {
          const int cb = m * 16 + 4 * g;
          if (cb >= a.ycp) continue;
          const long long o = pos * a.ycp + cb;
          u16 yv[4], zv4[4];
          u16x4 aux4 = {0, 0, 0, 0};
          if (EPI == CONV_EPI_GELU_BWD) aux4 = *reinterpret_cast<const u16x4*>(a.aux + o);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int co = cb + v;
            const bool real = co < a.cout;
            const float val = acc[m][n][v];
            if (EPI == CONV_EPI_GELU) {
              const float zv = real ? val + (a.bias ? a.bias[co] : 0.f) : 0.f;
              const u16 zb = f2bf(zv);
              zv4[v] = zb;
              yv[v] = real ? f2bf(gelu(bf2f(zb))) : (u16)0;
            } else if (EPI == CONV_EPI_LINEAR_BF16) {
              yv[v] = real ? f2bf(val + (a.bias ? a.bias[co] : 0.f)) : (u16)0;
            } else {  // CONV_EPI_GELU_BWD
              yv[v] = real ? f2bf(val * gelu_grad(bf2f(aux4[v]))) : (u16)0;
            }
          }
          *reinterpret_cast<u16x4*>(reinterpret_cast<u16*>(a.y) + o) =
              u16x4{yv[0], yv[1], yv[2], yv[3]};
          if (EPI == CONV_EPI_GELU && a.zout)
            *reinterpret_cast<u16x4*>(a.zout + o) = u16x4{zv4[0], zv4[1], zv4[2], zv4[3]};
          continue;
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int co = m * 16 + 4 * g + v;
          if (co >= a.ycp) continue;
          const bool real = co < a.cout;
          float val = acc[m][n][v];
          if (EPI == CONV_EPI_GELU) {
            const float zv = real ? val + (a.bias ? a.bias[co] : 0.f) : 0.f;
            const u16 zb = f2bf(zv);
            if (a.zout) a.zout[pos * a.ycp + co] = zb;
            // gelu of the bf16-rounded pre-activation (what backward sees)
            reinterpret_cast<u16*>(a.y)[pos * a.ycp + co] = real ? f2bf(gelu(bf2f(zb))) : (u16)0;
          } else if (EPI == CONV_EPI_LINEAR_F32) {
            if (real) reinterpret_cast<float*>(a.y)[pos * a.ycp + co] =
                val + (a.bias ? a.bias[co] : 0.f);
          } else if (EPI == CONV_EPI_LINEAR_BF16) {
            reinterpret_cast<u16*>(a.y)[pos * a.ycp + co] =
                real ? f2bf(val + (a.bias ? a.bias[co] : 0.f)) : (u16)0;
          } else {  // CONV_EPI_GELU_BWD
            const float zv = bf2f(a.aux[pos * a.ycp + co]);
            reinterpret_cast<u16*>(a.y)[pos * a.ycp + co] = real ? f2bf(val * gelu_grad(zv)) : (u16)0;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Row kernel for 8 or 32 input channels (every layer of the closure net and
// its input gradients).  Same GEMM as conv_fwd_kernel, arranged so that the
// K loop issues little besides loads, LDS reads and MFMAs:
//  * input rows are read through buffer resources whose size is the row's
//    bytes (zero for a row outside the grid): the hardware returns zero for
//    an offset past the end, so x-1 < 0, x >= nx and missing rows need no
//    per-lane test, and every address is a wave-uniform base plus a lane
//    offset fixed for the whole launch;
//  * the 9 (dz, dy) pairs run two at a time, the next pair's fragments (after
//    a plane's last pair, the next plane's first) in flight during the
//    current pair's MFMAs and the epilogue;
//  * (32 channels) one edge fragment per pair holds column x0 - 1 in lane 0
//    and x0 + 64 in lane 15 of each row group, so the outer tiles' shifts
//    take one DPP move per dword.
// The column tiles are 16 consecutive x each: a tile's load is one 1 KiB
// contiguous span.  (Interleaving the tiles, x = x0 + 4 r + n, would make the
// x-1 / x+1 fragments of inner tiles plain register copies, but spreads each
// load over 4 KiB: layer 3 of the net took 171 instead of 115 us.)
// 32 channels: one K step per tap, K order tap * 32 + channel as in the
// global weights.  8 channels: one K step per (dz, dy) pair, k group g of
// the B operand holding tap dx = g - 1 (g = 3 a zero column of the LDS
// weights), so each lane loads one fragment per tile and nothing is shifted.
// The bf16 GELU epilogue reads gelu of the bf16 pre-activation from a table
// in LDS for |z| in [2^-14, 8) (bf16 exponents 113..129, both signs), built
// by each block with the same gelu() as the direct path; the rare rest takes
// the direct path.
// Results are those of conv_fwd_kernel bit for bit with 32 channels; with 8
// channels the K steps group the taps differently (float32 accumulation
// order).
// ---------------------------------------------------------------------------
constexpr int kRowE0 = 113;            // smallest bf16 exponent in the table
constexpr int kRowLutN = 17 * 128;     // exponents 113..129, one sign
constexpr unsigned kRowWord3 = 0x00020000u;  // raw buffer, 32-bit format

__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const void* base, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes,
                                           (int)kRowWord3);
}
__device__ __forceinline__ bf16x8 row_ld(__amdgpu_buffer_rsrc_t rs, unsigned off) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)off, 0, 0);
  return __builtin_bit_cast(bf16x8, v);
}
// two floats -> two bf16 (round to nearest even, v_cvt_pk_bf16_f32), the
// first in the low half.  The empty asm keeps the packed word opaque, so
// that the table index below is computed from it (two values per packed
// 16-bit instruction) rather than rebuilt from the floats' sign bits.
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned cvt2(float lo, float hi) {
  const f32x2 v = {lo, hi};
  unsigned w = __builtin_bit_cast(unsigned, __builtin_convertvector(v, bf16x2));
  asm volatile("" : "+v"(w));
  return w;
}

// gelu of W words of two bf16 pre-activations each: t = (|z| bits) -
// 113 * 128 (mod 2^16) indexes the table when below kRowLutN (one unused
// entry per sign follows each half, so the index min(t, kRowLutN) needs no
// select); every other value (|z| >= 8, tiny, inf, nan: t >= kRowLutN)
// takes the direct path, one branch for all W words, rarely taken
constexpr int kRowLutS = kRowLutN + 1;  // table entries per sign
template <int W>
__device__ __forceinline__ void gelu_words(const u16* lut, const unsigned* zw, unsigned* yw) {
  u16x2 tt[W];
  unsigned lo[W], hi[W];
  u16x2 tmax = {0, 0};
#pragma unroll
  for (int w = 0; w < W; ++w) {
    const u16x2 z = __builtin_bit_cast(u16x2, zw[w]);
    const u16x2 tv = (z & (u16)0x7FFF) - (u16)(kRowE0 * 128);
    const u16x2 id =
        (__builtin_elementwise_min(tv, (u16x2)(u16)kRowLutN) + (z >> 15) * (u16)kRowLutS) << 1;
    tt[w] = tv;
    tmax = __builtin_elementwise_max(tmax, tv);
    const unsigned idw = __builtin_bit_cast(unsigned, id);
    lo[w] = idw & 0xFFFFu;
    hi[w] = idw >> 16;
  }
  const unsigned char* lb = reinterpret_cast<const unsigned char*>(lut);
#pragma unroll
  for (int w = 0; w < W; ++w)
    yw[w] = (unsigned)*reinterpret_cast<const u16*>(lb + lo[w]) |
            ((unsigned)*reinterpret_cast<const u16*>(lb + hi[w]) << 16);
  if (__builtin_elementwise_max(tmax.x, tmax.y) >= (u16)kRowLutN) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if (tt[w].x >= (u16)kRowLutN)
        yw[w] = (yw[w] & 0xFFFF0000u) | f2bf(gelu(bf2f((u16)(zw[w] & 0xFFFFu))));
      if (tt[w].y >= (u16)kRowLutN)
        yw[w] = (yw[w] & 0xFFFFu) | ((unsigned)f2bf(gelu(bf2f((u16)(zw[w] >> 16)))) << 16);
    }
  }
}

// bf16 outputs of one lane's 4 channels cb..cb+3, two per word (one 8-byte
// store each for y and, when given, the pre-activation z); channels past
// cout are written as zero (no test when every channel of the layer is real)
__device__ __forceinline__ void bf16_store4(const unsigned* yw, const unsigned* zw, int cb, int cout,
                                            bool allreal, __amdgpu_buffer_rsrc_t ors, const u16* zout,
                                            long long orow, unsigned obytes, unsigned o) {
  unsigned y0 = yw[0], y1 = yw[1], z0 = zw[0], z1 = zw[1];
  if (!allreal) {
    const unsigned m0 = (cb < cout ? 0xFFFFu : 0u) | (cb + 1 < cout ? 0xFFFF0000u : 0u);
    const unsigned m1 = (cb + 2 < cout ? 0xFFFFu : 0u) | (cb + 3 < cout ? 0xFFFF0000u : 0u);
    y0 &= m0; z0 &= m0; y1 &= m1; z1 &= m1;
  }
  const i32x2 yy = {(int)y0, (int)y1};
  __builtin_amdgcn_raw_buffer_store_b64(yy, ors, (int)o, 0, 0);
  if (zout) {
    const i32x2 zz = {(int)z0, (int)z1};
    __builtin_amdgcn_raw_buffer_store_b64(zz, row_rsrc(zout + orow, obytes), (int)o, 0, 0);
  }
}

// bf16 epilogue of one column tile: an[m] holds the sums of channels
// m * 16 + 4 g .. + 3 at position px
template <int EPI, int MT>
__device__ __forceinline__ void bf16_epi(const f32x4* an, const float (*bv)[4], const u16* lut,
                                         const ConvArgs& a, int g, int px, int ycp, bool allreal,
                                         __amdgpu_buffer_rsrc_t ors, long long orow, unsigned obytes) {
  unsigned zw[MT * 2] = {}, yw[MT * 2] = {};
  if (EPI == CONV_EPI_GELU || EPI == CONV_EPI_LINEAR_BF16) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const unsigned w = cvt2(an[m][2 * h] + bv[m][2 * h], an[m][2 * h + 1] + bv[m][2 * h + 1]);
        if (EPI == CONV_EPI_GELU) zw[m * 2 + h] = w;
        else yw[m * 2 + h] = w;
      }
  }
  if (EPI == CONV_EPI_GELU) gelu_words<MT * 2>(lut, zw, yw);
  if (EPI == CONV_EPI_GELU_BWD) {
#pragma unroll
    for (int m = 0; m < MT; ++m) {
      const unsigned o = (unsigned)(px * ycp + m * 16 + 4 * g) * 2u;
      const auto av = __builtin_amdgcn_raw_buffer_load_b64(row_rsrc(a.aux + orow, obytes), (int)o, 0, 0);
      const u16x4 aux4 = __builtin_bit_cast(u16x4, av);
#pragma unroll
      for (int h = 0; h < 2; ++h)
        yw[m * 2 + h] = (unsigned)f2bf(an[m][2 * h] * gelu_grad(bf2f(aux4[2 * h]))) |
                        ((unsigned)f2bf(an[m][2 * h + 1] * gelu_grad(bf2f(aux4[2 * h + 1]))) << 16);
    }
  }
#pragma unroll
  for (int m = 0; m < MT; ++m) {
    const int cb = m * 16 + 4 * g;
    if (cb >= ycp) continue;
    const unsigned o = (unsigned)(px * ycp + cb) * 2u;
    bf16_store4(yw + m * 2, zw + m * 2, cb, a.cout, allreal, ors,
                EPI == CONV_EPI_GELU ? a.zout : nullptr, orow, obytes, o);
  }
}

// float32 output of one lane's 4 channels cb..cb+3 at position px: one
// 16-byte store when the channel stride is a multiple of 4 (channels past
// cout written as zero), else one 4-byte store per real channel
typedef int i32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void f32_out(const f32x4& acc, const float* bv, int cb, int cout, int ycp,
                                        __amdgpu_buffer_rsrc_t ors, int px) {
  if ((ycp & 3) == 0) {
    if (cb >= ycp) return;
    f32x4 o;
#pragma unroll
    for (int v = 0; v < 4; ++v) o[v] = cb + v < cout ? acc[v] + bv[v] : 0.f;
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(i32x4, o), ors,
                                           (int)((unsigned)(px * ycp + cb) * 4u), 0, 0);
    return;
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int co = cb + v;
    if (co < cout && co < ycp)
      __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, acc[v] + bv[v]), ors,
                                            (int)((unsigned)(px * ycp + co) * 4u), 0, 0);
  }
}

template <int EPI, int MT, int CIN>
__global__ __launch_bounds__(1024) void conv_row_kernel(ConvArgs a) {
  constexpr int NT = 4;
  constexpr int KP = CIN == 32 ? 27 * 32 : 9 * 32;  // K per output channel in LDS
  constexpr int LDW = KP + 8;
  extern __shared__ u16 sh[];
  u16* wsh = sh;                          // [MT * 16][LDW]
  u16* lut = sh + MT * 16 * LDW;          // [2][kRowLutS] (GELU only)
  {
    constexpr int C16 = KP / 8;           // 16-byte chunks per output channel
    for (int q = threadIdx.x; q < MT * 16 * C16; q += 1024) {
      const int row = q / C16, c = q - row * C16;
      uint4 v = {0u, 0u, 0u, 0u};
      if (CIN == 32) {
        v = *reinterpret_cast<const uint4*>(a.w + (long long)row * a.kpad + c * 8);
      } else {
        const int p = c >> 2, gg = c & 3;
        if (gg < 3) v = *reinterpret_cast<const uint4*>(a.w + (long long)row * a.kpad + (3 * p + gg) * 8);
      }
      *reinterpret_cast<uint4*>(wsh + row * LDW + c * 8) = v;
    }
    if (EPI == CONV_EPI_GELU) {
      for (int q = threadIdx.x; q < 2 * kRowLutS; q += 1024) {
        const unsigned neg = q >= kRowLutS, j = q - neg * kRowLutS;
        const unsigned zb = (neg << 15) | (kRowE0 * 128u + j);
        lut[q] = j < (unsigned)kRowLutN ? f2bf(gelu(bf2f((u16)zb))) : (u16)0;
      }
    }
  }
  __syncthreads();
  const int t = threadIdx.x & 63;
  const int g = t >> 4, r = t & 15;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int yy = blockIdx.y * 16 + wv;
  if (yy >= ny) return;
  const int x0 = blockIdx.x * 16 * NT;
  const int z0 = blockIdx.z * a.zchunk;
  const int z1 = min(z0 + a.zchunk, nz);
  const unsigned rowbytes = (unsigned)nx * CIN * 2u;
  // lane offsets within an input row (bytes; negative wraps past the end)
  // (32 channels) the edge fragment: column x0 - 1 in lane 0 of each row
  // group (the x-1 tap of tile 0), x0 + 64 in lane 15 (x+1 of tile 3)
  unsigned off[NT], offe = 0;
  if (CIN == 32) {
#pragma unroll
    for (int n = 0; n < NT; ++n) off[n] = (unsigned)((x0 + 16 * n + r) * 64 + 16 * g);
    offe = (unsigned)((r == 15 ? x0 + 16 * NT : x0 - 1) * 64 + 16 * g);
  } else {
    const int dx = g < 3 ? g - 1 : 0;     // g = 3 meets zero weights
#pragma unroll
    for (int n = 0; n < NT; ++n) off[n] = (unsigned)((x0 + 16 * n + r + dx) * 16);
  }
  const int ycp = a.ycp;
  const bool allreal = a.cout >= MT * 16;
  float bv[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int co = m * 16 + 4 * g + v;
      bv[m][v] = a.bias && co < a.cout ? a.bias[co] : 0.f;
    }
  const unsigned esz = EPI == CONV_EPI_LINEAR_F32 ? 4u : 2u;
  const unsigned obytes = (unsigned)nx * ycp * esz;
  auto rsrc_of = [&](int zz, int p) {
    const int sz = zz + p / 3 - 1, sy = yy + p % 3 - 1;
    const bool pl = sz >= 0 && sz < nz && sy >= 0 && sy < ny;
    const u16* base = a.x + ((long long)(pl ? sz : 0) * ny + (pl ? sy : 0)) * nx * CIN;
    return row_rsrc(base, pl ? rowbytes : 0u);
  };
  // fragments of one (dz, dy) pair: the NT x+0 tiles and the edge fragment
  struct Pair {
    bf16x8 c[NT], e;
  };
  auto load_pair = [&](int zz, int p, Pair& f) {
    const auto rs = rsrc_of(zz, p);
#pragma unroll
    for (int n = 0; n < NT; ++n) f.c[n] = row_ld(rs, off[n]);
    if (CIN == 32) f.e = row_ld(rs, offe);
  };
  // two fragment sets: the next pair's loads (after the last pair of a
  // plane, the next plane's first) are in flight during the current pair's
  // MFMAs and the epilogue
  Pair fa, fb;
  load_pair(z0, 0, fa);
  for (int zz = z0; zz < z1; ++zz) {
    f32x4 acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    auto mma_pair = [&](int p, const Pair& f) {
      if (CIN == 32) {
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          bf16x8 b[NT];
#pragma unroll
          for (int n = 0; n < NT; ++n) {
            if (dx == 0) b[n] = n == 0 ? shift_in_right(f.c[0], f.e) : shift_right(f.c[n], f.c[n - 1]);
            else if (dx == 1) b[n] = f.c[n];
            else b[n] = n == NT - 1 ? shift_in_left(f.c[NT - 1], f.e) : shift_left(f.c[n], f.c[n + 1]);
          }
          bf16x8 af[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) af[m] = load8(wsh + (m * 16 + r) * LDW + (3 * p + dx) * 32 + 8 * g);
#pragma unroll
          for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int m = 0; m < MT; ++m)
              acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], b[n], acc[m][n], 0, 0, 0);
        }
      } else {
        bf16x8 af[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) af[m] = load8(wsh + (m * 16 + r) * LDW + p * 32 + 8 * g);
#pragma unroll
        for (int n = 0; n < NT; ++n)
#pragma unroll
          for (int m = 0; m < MT; ++m)
            acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], f.c[n], acc[m][n], 0, 0, 0);
      }
    };
    // fa holds pair p on entry to each iteration
#pragma unroll 1
    for (int p = 0; p < 9; p += 2) {
      if (p + 1 < 9) load_pair(zz, p + 1, fb);
      else if (zz + 1 < z1) load_pair(zz + 1, 0, fb);
      mma_pair(p, fa);
      if (p + 1 < 9) {
        if (p + 2 < 9) load_pair(zz, p + 2, fa);
        mma_pair(p + 1, fb);
      }
    }
    // epilogue: lane holds D[co = m*16 + 4g + v][x = x0 + 16n + r]; stores past
    // the row's end (x >= nx) fall outside the buffer and are dropped
    const long long orow = ((long long)zz * ny + yy) * nx * ycp;
    char* ybase = reinterpret_cast<char*>(a.y) + orow * esz;
    const auto ors = row_rsrc(ybase, obytes);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int px = x0 + 16 * n + r;
#pragma unroll
      for (int m = 0; m < MT; ++m) {
        const int cb = m * 16 + 4 * g;
        if (EPI == CONV_EPI_LINEAR_F32) {
          f32_out(acc[m][n], bv[m], cb, a.cout, ycp, ors, px);
          continue;
        }
      }
      if (EPI == CONV_EPI_LINEAR_F32) continue;
      // bf16 outputs: the lane's 4 consecutive channels per M tile in one
      // 8-byte store; channels past cout are written as zero
      {
        f32x4 an[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) an[m] = acc[m][n];
        bf16_epi<EPI, MT>(an, bv, lut, a, g, px, ycp, allreal, ors, orow, obytes);
      }
    }
    fa = fb;
  }
}

// ---------------------------------------------------------------------------
// Slab kernel for 32 input channels: the input planes a block needs are
// staged in LDS by direct-to-LDS buffer loads and every wave reads its
// fragments there, so each input row is fetched from the caches about 2.6
// times per output plane instead of 9.
//  * a block (16 wavefronts) computes 64 x by 8 rows by 2 planes per step;
//    wave w owns row y0 + (w & 7) of plane zz + (w >> 3); lane column r of
//    column tile n is x = x0 + 4 r + n, so the x-1 / x+1 fragments of tile n
//    are the x+0 fragments of tiles n-1 / n+1 and only tiles 0 and 3 need a
//    DPP shift (with an edge fragment: x0 - 1 in lane 0, x0 + 64 in lane 15);
//  * a slab is one input plane's rows y0-1 .. y0+8 and columns x0-1 .. x0+64;
//    within a row, position p's channel chunk g (16 bytes) sits at slot
//    17 (4 (p & 3) + g) + (p >> 2): the 16 lanes of a row group, which read
//    positions 4 r + c, read 16 consecutive slots, the one pattern in which
//    ds_read_b128 runs at the full LDS rate on gfx950 (a 16-lane group
//    spread over more than 256 bytes runs at half of it or less);
//  * the weights sit in LDS as [tap][k chunk g][output channel r] 16-byte
//    slots for the same reason;
//  * each lane of a direct-to-LDS load writes 16 bytes at consecutive LDS
//    addresses, so every slot's global source is chosen per lane; positions
//    and rows outside the grid (and the unused slots) get an offset past the
//    buffer's end and arrive as zeros, a plane outside the grid a buffer of
//    size zero;
//  * the four slabs iz = zz-1 .. zz+2 of a step stream through two LDS
//    buffers, the next slab's loads in flight during the current slab's
//    MFMAs; plane zz's waves have nothing to add from slab zz+2 and plane
//    zz+1's nothing from slab zz-1: each wave writes its finished row in
//    that idle slab step (plane zz+1's at the next step's first slab).
// Taps accumulate in the order (dz, dy, dx) of conv_fwd_kernel, so the
// results are the same bits.
// PX (float32 outputs, at most 8 channels: the closure net's last layer)
// stacks two taps in the 16 output rows of one MFMA instead of padding 6
// channels to 16: per (dz, dy) the fragment of input column x meets rows
// 0-7 = tap dx = -1 (a term of output x+1) and rows 8-15 = tap dx = +1 (of
// output x-1) in one MFMA, tap dx = 0 in a second, and the edge fragment
// (x0 - 1, x0 + 64) the first once more: 9 MFMAs per (dz, dy) and wave
// instead of 12, and no shifted fragments.  The epilogue moves the x-1
// parts from lane groups 2-3 to 0-1 (ds_bpermute), shifts the x+1 / x-1
// parts one column (DPP, the edge columns filling lanes 0 and 15) and adds
// the three sums: float32 results within rounding of the taps' sum, not the
// bits of the (dz, dy, dx) order.
// ---------------------------------------------------------------------------
constexpr int kSlabRows = 10, kSlabRowSlots = 272;          // 66 positions (+ 16 unused slots)
constexpr int kSlabSlots = kSlabRows * kSlabRowSlots;        // 16-byte slots
constexpr int kSlabBytes = kSlabSlots * 16;                  // 43520
constexpr int kSlabLoads = (kSlabSlots + 63) / 64;           // 43 wave loads per slab

__device__ __forceinline__ int slab_slot(int pos, int g) { return 17 * (4 * (pos & 3) + g) + (pos >> 2); }

template <int EPI, int MT, bool PX = false>
__global__ __launch_bounds__(1024) void conv_slab_kernel(ConvArgs a) {
  constexpr int NT = 4, CIN = 32;
  static_assert(!PX || (MT == 1 && EPI == CONV_EPI_LINEAR_F32), "paired taps: f32, one M tile");
  // separate LDS objects (one per weight M tile, one per slab buffer), so
  // that the compiler can tell the slab being read and the weights from the
  // slab being loaded and does not wait for those loads
  __shared__ __attribute__((aligned(16))) uint4 wsh0[PX ? 9 * 2 * 4 * 16 : 27 * 4 * 16];
  __shared__ __attribute__((aligned(16))) uint4 wsh1[MT == 2 ? 27 * 4 * 16 : 1];
  __shared__ u16 lut[EPI == CONV_EPI_GELU ? 2 * kRowLutS : 2];
  __shared__ __attribute__((aligned(16))) unsigned char slab0[kSlabBytes];
  __shared__ __attribute__((aligned(16))) unsigned char slab1[kSlabBytes];
  if constexpr (PX) {
    // slot ((2 p + h) 4 + g) 16 + row for (dz, dy) pair p = 3 (dz + 1) +
    // dy: h = 0 holds tap dx = -1 in rows 0-7 (output channel row) and
    // dx = +1 in rows 8-15 (channel row - 8), h = 1 tap dx = 0 in rows 0-7
    // and zeros; k chunk g = channels 8 g .. 8 g + 7
    for (int q = threadIdx.x; q < 9 * 2 * 4 * 16; q += 1024) {
      const int row = q & 15, c = q >> 4;
      const int gc = c & 3, h = (c >> 2) & 1, p = c >> 3;
      const int co = row & 7, dx = h ? 1 : (row < 8 ? 0 : 2);
      const bool ok = co < a.cout && !(h && row >= 8);
      wsh0[q] = ok ? *reinterpret_cast<const uint4*>(a.w + (long long)co * a.kpad +
                                                     (p * 3 + dx) * 32 + gc * 8)
                   : make_uint4(0u, 0u, 0u, 0u);
    }
  } else {
    // k chunk c = 4 tap + g (8 channels) of output channel row -> slot (c, row)
    constexpr int C16 = 27 * 32 / 8;
    for (int q = threadIdx.x; q < MT * 16 * C16; q += 1024) {
      const int row = q / C16, c = q - row * C16;
      (row < 16 ? wsh0 : wsh1)[c * 16 + (row & 15)] =
          *reinterpret_cast<const uint4*>(a.w + (long long)row * a.kpad + c * 8);
    }
  }
  {
    if (EPI == CONV_EPI_GELU) {
      for (int q = threadIdx.x; q < 2 * kRowLutS; q += 1024) {
        const unsigned neg = q >= kRowLutS, j = q - neg * kRowLutS;
        const unsigned zb = (neg << 15) | (kRowE0 * 128u + j);
        lut[q] = j < (unsigned)kRowLutN ? f2bf(gelu(bf2f((u16)zb))) : (u16)0;
      }
    }
  }
  const int t = threadIdx.x & 63;
  const int g = t >> 4, r = t & 15;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 8;
  const int z0 = blockIdx.z * a.zchunk, z1 = min(z0 + a.zchunk, nz);
  const int jw = wv >> 3, iw = wv & 7;
  const int yy = y0 + iw;
  const bool rowok = yy < ny;
  // this lane's slots of the block's slab loads (instruction k = wv + 16 u)
  unsigned soff[3];
#pragma unroll
  for (int u = 0; u < 3; ++u) {
    const int q = (wv + 16 * u) * 64 + t;
    const int row = q / kSlabRowSlots, rs = q - row * kSlabRowSlots;
    const int cg = rs / 17, qd = rs - 17 * cg;
    const int pos = 4 * qd + (cg >> 2), gg = cg & 3;
    const int sy = y0 - 1 + row, sx = x0 - 1 + pos;
    const bool ok = q < kSlabSlots && pos < 66 && sy >= 0 && sy < ny && sx >= 0 && sx < nx;
    soff[u] = ok ? (unsigned)(((long long)sy * nx + sx) * 64 + 16 * gg) : 0x80000000u;
  }
  const unsigned planebytes = (unsigned)((long long)nx * ny * CIN * 2);
  auto stage = [&](int iz, unsigned char* dst) __attribute__((always_inline)) {
    const bool pl = iz >= 0 && iz < nz;
    const auto rs = row_rsrc(a.x + (long long)(pl ? iz : 0) * ny * nx * CIN, pl ? planebytes : 0u);
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      const int k = wv + 16 * u;
      if (k < kSlabLoads && (k * 64 + t) < kSlabSlots)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(dst + k * 1024), 16, soff[u], 0, 0, 0);
    }
  };
  const int ycp = a.ycp;
  const bool allreal = a.cout >= MT * 16;
  float bv[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int co = m * 16 + 4 * g + v;
      bv[m][v] = a.bias && co < a.cout ? a.bias[co] : 0.f;
    }
  const unsigned esz = EPI == CONV_EPI_LINEAR_F32 ? 4u : 2u;
  const unsigned obytes = (unsigned)nx * ycp * esz;
  // LDS byte offsets of this lane's fragments within a slab row
  unsigned loff[NT], eoff;
#pragma unroll
  for (int n = 0; n < NT; ++n) loff[n] = slab_slot(4 * r + n + 1, g) * 16;
  eoff = slab_slot(r == 15 ? 65 : 0, g) * 16;
  const int nsteps = (z1 - z0 + 1) / 2;
  const int nslab = 4 * nsteps;
  f32x4 acc[MT][NT];
  // PX: acc holds the dx = -1 / +1 pair, acc2 the dx = 0 taps, acce the
  // edge fragment's pair
  f32x4 acc2[PX ? NT : 1], acce;
  const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int n = 0; n < NT; ++n) acc[m][n] = zero4;
#pragma unroll
  for (int n = 0; n < (PX ? NT : 1); ++n) acc2[n] = zero4;
  acce = zero4;
  // write this wave's finished row of plane zw, then clear the sums
  auto epilogue = [&](int zw) __attribute__((always_inline)) {
    if (PX && rowok && zw < z1) {
      const long long orow = ((long long)zw * ny + yy) * nx * ycp;
      const auto ors = row_rsrc(reinterpret_cast<char*>(a.y) + orow * 4, obytes);
      // the x-1 parts (rows 8-15, lane groups 2-3) to lane groups 0-1
      f32x4 mv[NT], me;
      const int src = (t ^ 32) << 2;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int n = 0; n < NT; ++n)
          mv[n][v] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(acc[0][n][v])));
        me[v] = __int_as_float(__builtin_amdgcn_ds_bpermute(src, __float_as_int(acce[v])));
      }
      // output x0 + 4 r + n: acc2 of tile n, the x+1 part of column x - 1
      // (tile n - 1; for n = 0 tile 3 of lane r - 1, the edge x0 - 1 in
      // lane 0) and the x-1 part of column x + 1 (tile n + 1; for n = 3
      // tile 0 of lane r + 1, the edge x0 + 64 in lane 15)
      const f32x4 pl = __builtin_bit_cast(
          f32x4, shift_in_right(__builtin_bit_cast(bf16x8, acc[0][NT - 1]), __builtin_bit_cast(bf16x8, acce)));
      const f32x4 mr = __builtin_bit_cast(
          f32x4, shift_in_left(__builtin_bit_cast(bf16x8, mv[0]), __builtin_bit_cast(bf16x8, me)));
      if (g < 2) {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          const f32x4 o = acc2[n] + (n == 0 ? pl : acc[0][n - 1]) + (n == NT - 1 ? mr : mv[n + 1]);
          f32_out(o, bv[0], 4 * g, a.cout, ycp, ors, x0 + 4 * r + n);
        }
      }
    }
    if (!PX && rowok && zw < z1) {
      const long long orow = ((long long)zw * ny + yy) * nx * ycp;
      char* ybase = reinterpret_cast<char*>(a.y) + orow * esz;
      const auto ors = row_rsrc(ybase, obytes);
#pragma unroll
      for (int n = 0; n < NT; ++n) {
        const int px = x0 + 4 * r + n;
        if (EPI == CONV_EPI_LINEAR_F32) {
#pragma unroll
          for (int m = 0; m < MT; ++m) f32_out(acc[m][n], bv[m], m * 16 + 4 * g, a.cout, ycp, ors, px);
          continue;
        }
        {
          f32x4 an[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m) an[m] = acc[m][n];
          bf16_epi<EPI, MT>(an, bv, lut, a, g, px, ycp, allreal, ors, orow, obytes);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = zero4;
#pragma unroll
    for (int n = 0; n < (PX ? NT : 1); ++n) acc2[n] = zero4;
    acce = zero4;
  };
  // slab ts is read from sb while slab ts + 1 loads into nxt
  auto slab_step = [&](int ts, const unsigned char* sb, unsigned char* nxt)
                       __attribute__((always_inline)) {
    const int zz = z0 + 2 * (ts >> 2);
    const int q = ts & 3;
    const int iz = zz - 1 + q;
    if (ts + 1 < nslab) stage(z0 + 2 * ((ts + 1) >> 2) - 1 + ((ts + 1) & 3), nxt);
    if (jw == 0 && q == 3) epilogue(zz);
    if (jw == 1 && q == 0 && ts > 0) epilogue(zz - 1);
    const int zw = zz + jw;
    const int dz = iz - zw;
    if (PX && rowok && zw < z1 && dz >= -1 && dz <= 1) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const unsigned char* rowp = sb + (iw + dy) * (kSlabRowSlots * 16);
        bf16x8 c[NT], e;
#pragma unroll
        for (int n = 0; n < NT; ++n) c[n] = *reinterpret_cast<const bf16x8*>(rowp + loff[n]);
        e = *reinterpret_cast<const bf16x8*>(rowp + eoff);
        const int p = (dz + 1) * 3 + dy;
        const bf16x8 a1 = __builtin_bit_cast(bf16x8, wsh0[((2 * p) * 4 + g) * 16 + r]);
        const bf16x8 a2 = __builtin_bit_cast(bf16x8, wsh0[((2 * p + 1) * 4 + g) * 16 + r]);
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          acc[0][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, c[n], acc[0][n], 0, 0, 0);
          acc2[n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, c[n], acc2[n], 0, 0, 0);
        }
        acce = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, e, acce, 0, 0, 0);
      }
    }
    if (!PX && rowok && zw < z1 && dz >= -1 && dz <= 1) {
#pragma unroll
      for (int dy = 0; dy < 3; ++dy) {
        const unsigned char* rowp = sb + (iw + dy) * (kSlabRowSlots * 16);
        bf16x8 c[NT], e;
#pragma unroll
        for (int n = 0; n < NT; ++n) c[n] = *reinterpret_cast<const bf16x8*>(rowp + loff[n]);
        e = *reinterpret_cast<const bf16x8*>(rowp + eoff);
        const int tap0 = ((dz + 1) * 3 + dy) * 3;
        const bf16x8 lft = shift_in_right(c[NT - 1], e), rgt = shift_in_left(c[0], e);
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
          bf16x8 b[NT];
#pragma unroll
          for (int n = 0; n < NT; ++n) {
            if (dx == 0) b[n] = n == 0 ? lft : c[n - 1];
            else if (dx == 1) b[n] = c[n];
            else b[n] = n == NT - 1 ? rgt : c[n + 1];
          }
          bf16x8 af[MT];
#pragma unroll
          for (int m = 0; m < MT; ++m)
            af[m] = __builtin_bit_cast(bf16x8, (m == 0 ? wsh0 : wsh1)[((tap0 + dx) * 4 + g) * 16 + r]);
#pragma unroll
          for (int n = 0; n < NT; ++n)
#pragma unroll
            for (int m = 0; m < MT; ++m)
              acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], b[n], acc[m][n], 0, 0, 0);
        }
      }
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's slab loads
    __syncthreads();
  };
  stage(z0 - 1, slab0);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
  for (int ts = 0; ts < nslab; ts += 2) {  // nslab is even
    slab_step(ts, slab0, slab1);
    slab_step(ts + 1, slab1, slab0);
  }
  if (jw == 1) epilogue(z0 + 2 * (nsteps - 1) + 1);
}

// ---------------------------------------------------------------------------
// Rolling-slab kernel for 8 input channels (the closure net's first layer,
// and input gradients into 8-channel tensors).  A block (16 wavefronts)
// computes 64 x by 16 rows, one plane per step; wave w owns row y0 + w.  An
// input plane's rows y0-1 .. y0+16 and columns x0-1 .. x0+64 (18 x 66
// positions of 16 bytes, 19 KiB) fill one of four LDS buffers: planes
// zz-1, zz, zz+1 are read while plane zz+2 loads (direct-to-LDS), so every
// input plane is fetched once per block.  K steps of four taps (k group g
// of the B operand one tap, read straight from the slab at the tap's row
// and x offset, no shifts): two per input plane for its in-plane taps 0-7,
// and one for the last in-plane tap of the three planes (lane group g reads
// plane g; group 3 meets zero weights) -- 7 steps where one per (dz, dy)
// pair with a zero fourth group would take 9.  (The row kernel groups the
// taps by (dz, dy): the float32 sums may differ in the last bit.)
// ---------------------------------------------------------------------------
constexpr int kRollRows = 18, kRollPos = 66;
constexpr int kRollSlots = kRollRows * kRollPos;             // 16-byte slots
constexpr int kRollBytes = kRollSlots * 16;                  // 19008
constexpr int kRollLoads = (kRollSlots + 63) / 64;           // 19 wave loads per plane

// tap of K step s, k group g in the 8-channel kernel: steps 0-5 take the 9
// taps of plane s / 2 four at a time (in-plane taps 0-3, 4-7), step 6 the
// last in-plane tap of the three planes; 27 (no tap) for the spare group
__host__ __device__ constexpr int roll8_tap(int s, int g) {
  return s < 6 ? (s >> 1) * 9 + 4 * (s & 1) + g : (g < 3 ? g * 9 + 8 : 27);
}

template <int EPI, int MT>
__global__ __launch_bounds__(1024) void conv_roll8_kernel(ConvArgs a) {
  constexpr int NT = 4, CIN = 8;
  // separate LDS objects, so that the compiler can tell each buffer being
  // read from the one being loaded; weights as [pair][k group g][output
  // channel r] 16-byte slots (a row group reads 16 consecutive slots)
  __shared__ __attribute__((aligned(16))) uint4 wsh0[7 * 4 * 16];
  __shared__ __attribute__((aligned(16))) uint4 wsh1[MT == 2 ? 7 * 4 * 16 : 1];
  __shared__ u16 lut[EPI == CONV_EPI_GELU ? 2 * kRowLutS : 2];
  __shared__ __attribute__((aligned(16))) unsigned char sl0[kRollBytes];
  __shared__ __attribute__((aligned(16))) unsigned char sl1[kRollBytes];
  __shared__ __attribute__((aligned(16))) unsigned char sl2[kRollBytes];
  __shared__ __attribute__((aligned(16))) unsigned char sl3[kRollBytes];
  {
    // K step s, k group g -> tap roll8_tap(s, g) (8 channels); 27 taps in
    // 7 steps, the last slot zero
    constexpr int C16 = 7 * 4;
    for (int q = threadIdx.x; q < MT * 16 * C16; q += 1024) {
      const int row = q / C16, c = q - row * C16;
      const int tap = roll8_tap(c >> 2, c & 3);
      uint4 v = {0u, 0u, 0u, 0u};
      if (tap < 27) v = *reinterpret_cast<const uint4*>(a.w + (long long)row * a.kpad + tap * 8);
      (row < 16 ? wsh0 : wsh1)[c * 16 + (row & 15)] = v;
    }
    if (EPI == CONV_EPI_GELU) {
      for (int q = threadIdx.x; q < 2 * kRowLutS; q += 1024) {
        const unsigned neg = q >= kRowLutS, j = q - neg * kRowLutS;
        const unsigned zb = (neg << 15) | (kRowE0 * 128u + j);
        lut[q] = j < (unsigned)kRowLutN ? f2bf(gelu(bf2f((u16)zb))) : (u16)0;
      }
    }
  }
  const int t = threadIdx.x & 63;
  const int g = t >> 4, r = t & 15;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int x0 = blockIdx.x * 64, y0 = blockIdx.y * 16;
  const int z0 = blockIdx.z * a.zchunk, z1 = min(z0 + a.zchunk, nz);
  const int yy = y0 + wv;
  const bool rowok = yy < ny;
  // this lane's slots of the block's plane loads (instruction k = wv + 16 u)
  unsigned soff[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int q = (wv + 16 * u) * 64 + t;
    const int row = q / kRollPos, pos = q - row * kRollPos;
    const int sy = y0 - 1 + row, sx = x0 - 1 + pos;
    const bool ok = q < kRollSlots && sy >= 0 && sy < ny && sx >= 0 && sx < nx;
    soff[u] = ok ? (unsigned)(((long long)sy * nx + sx) * 16) : 0x80000000u;
  }
  const unsigned planebytes = (unsigned)((long long)nx * ny * CIN * 2);
  auto stage = [&](int iz, unsigned char* dst) __attribute__((always_inline)) {
    const bool pl = iz >= 0 && iz < nz;
    const auto rs = row_rsrc(a.x + (long long)(pl ? iz : 0) * ny * nx * CIN, pl ? planebytes : 0u);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int k = wv + 16 * u;
      if (k < kRollLoads && (k * 64 + t) < kRollSlots)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(
            rs, (__attribute__((address_space(3))) void*)(dst + k * 1024), 16, soff[u], 0, 0, 0);
    }
  };
  float bv[MT][4];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int co = m * 16 + 4 * g + v;
      bv[m][v] = a.bias && co < a.cout ? a.bias[co] : 0.f;
    }
  const int ycp = a.ycp;
  const bool allreal = a.cout >= MT * 16;
  const unsigned esz = EPI == CONV_EPI_LINEAR_F32 ? 4u : 2u;
  const unsigned obytes = (unsigned)nx * ycp * esz;
  // this lane's byte offset in a plane buffer for tile 0 of the K steps
  // that read one plane (k = s % 2: in-plane tap 4 k + g, (dy, dx) =
  // (j / 3 - 1, j % 3 - 1)) and of the step that reads tap (1, 1) of all
  // three; tile n adds 256 n
  unsigned boff[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const int j = k < 2 ? 4 * k + g : 8;
    boff[k] = (unsigned)((wv + j / 3) * (kRollPos * 16) + (r + j % 3) * 16);
  }
  const int nplanes = z1 - z0;
  f32x4 acc[MT][NT];
  // fragments of one (dz, dy) pair; the next pair's six LDS reads are issued
  // before the current pair's MFMAs (the scheduling barrier keeps them there:
  // left alone, the compiler reads each fragment just before its two MFMAs
  // and waits for it)
  struct Frag {
    bf16x8 b[NT], af[MT];
  };
  auto compute = [&](const unsigned char* pm, const unsigned char* p0, const unsigned char* pp)
                     __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int n = 0; n < NT; ++n) acc[m][n] = f32x4{0.f, 0.f, 0.f, 0.f};
    // steps 0-5 read plane s / 2 (zz - 1, zz, zz + 1); step 6 reads plane g
    // in lane group g (group 3 meets zero weights: any valid position)
    auto load = [&](int p, Frag& f) __attribute__((always_inline)) {
      if (p < 6) {
        const unsigned char* sb = (p >> 1) == 0 ? pm : ((p >> 1) == 1 ? p0 : pp);
#pragma unroll
        for (int n = 0; n < NT; ++n)
          f.b[n] = *reinterpret_cast<const bf16x8*>(sb + boff[p & 1] + 256 * n);
      } else {
#pragma unroll
        for (int n = 0; n < NT; ++n) {
          if (g == 0) f.b[n] = *reinterpret_cast<const bf16x8*>(pm + boff[2] + 256 * n);
          else if (g == 1) f.b[n] = *reinterpret_cast<const bf16x8*>(p0 + boff[2] + 256 * n);
          else f.b[n] = *reinterpret_cast<const bf16x8*>(pp + boff[2] + 256 * n);
        }
      }
#pragma unroll
      for (int m = 0; m < MT; ++m)
        f.af[m] = __builtin_bit_cast(bf16x8, (m == 0 ? wsh0 : wsh1)[(p * 4 + g) * 16 + r]);
    };
    Frag f[2];
    load(0, f[0]);
#pragma unroll
    for (int p = 0; p < 7; ++p) {
      if (p + 1 < 7) load(p + 1, f[(p + 1) & 1]);
      __builtin_amdgcn_sched_barrier(0);
      const Frag& c = f[p & 1];
#pragma unroll
      for (int n = 0; n < NT; ++n)
#pragma unroll
        for (int m = 0; m < MT; ++m)
          acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(c.af[m], c.b[n], acc[m][n], 0, 0, 0);
    }
  };
  auto epilogue = [&](int zz) __attribute__((always_inline)) {
    if (!rowok) return;
    const long long orow = ((long long)zz * ny + yy) * nx * ycp;
    char* ybase = reinterpret_cast<char*>(a.y) + orow * esz;
    const auto ors = row_rsrc(ybase, obytes);
#pragma unroll
    for (int n = 0; n < NT; ++n) {
      const int px = x0 + 16 * n + r;
      if (EPI == CONV_EPI_LINEAR_F32) {
#pragma unroll
        for (int m = 0; m < MT; ++m) f32_out(acc[m][n], bv[m], m * 16 + 4 * g, a.cout, ycp, ors, px);
        continue;
      }
      {
        f32x4 an[MT];
#pragma unroll
        for (int m = 0; m < MT; ++m) an[m] = acc[m][n];
        bf16_epi<EPI, MT>(an, bv, lut, a, g, px, ycp, allreal, ors, orow, obytes);
      }
    }
  };
  // plane zz - 1 is written at the start of step zz: its stores then drain
  // during plane zz's MFMAs, and the wait for this wave's plane loads at the
  // end of the step (the same counter counts stores) finds them done
  auto plane_step = [&](int s, const unsigned char* pm, const unsigned char* p0, const unsigned char* pp,
                        unsigned char* nxt) __attribute__((always_inline)) {
    const int zz = z0 + s;
    if (s > 0) epilogue(zz - 1);
    if (s + 1 < nplanes) stage(zz + 2, nxt);
    if (rowok) compute(pm, p0, pp);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's plane loads
    __syncthreads();
  };
  stage(z0 - 1, sl0);
  stage(z0, sl1);
  stage(z0 + 1, sl2);
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
  __syncthreads();
  // plane z0 + s reads buffers s, s+1, s+2 (mod 4) and loads into s+3
  for (int s = 0; s < nplanes; s += 4) {
    plane_step(s, sl0, sl1, sl2, sl3);
    if (s + 1 < nplanes) plane_step(s + 1, sl1, sl2, sl3, sl0);
    if (s + 2 < nplanes) plane_step(s + 2, sl2, sl3, sl0, sl1);
    if (s + 3 < nplanes) plane_step(s + 3, sl3, sl0, sl1, sl2);
  }
  epilogue(z0 + nplanes - 1);
}

// ---------------------------------------------------------------------------
// weight gradient
// ---------------------------------------------------------------------------
struct WgradArgs {
  int nx, ny, nz, cin, cout_pad, dycp, kpad, nseg_x, segs_per_block, total_segs;
  const u16* x;
  const u16* dy;
  float* part;  // [gridDim.x][cout_pad][kpad]
};

constexpr int kSeg = 32;  // positions per reduction step (one MFMA K)

// Each block: 4 waves; wave w owns N tiles (k columns) n = w, w+4, ...;
// all waves share the M tiles (output channels).  LDS: dY^T [cout_pad][32]
// and X^T for the 9 neighbour rows [9][cin][kSeg + 2] of one segment.
template <int MT>
__global__ __launch_bounds__(256) void conv_wgrad_kernel(WgradArgs a) {
  extern __shared__ u16 sh[];
  u16* dyT = sh;                                   // [cout_pad][kSeg]
  u16* xT = sh + a.cout_pad * kSeg;                // [9][cin][kSeg + 2]
  const int xs = kSeg + 2;
  const int t = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int g = t >> 4, r = t & 15;
  const int ntile = a.kpad / 16;
  constexpr int kMaxTiles = 16;  // N tiles per wave (kpad <= 4*16*16 = 1024)
  f32x4 acc[MT][kMaxTiles];
#pragma unroll
  for (int m = 0; m < MT; ++m)
#pragma unroll
    for (int q = 0; q < kMaxTiles; ++q) acc[m][q] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int cin = a.cin, nx = a.nx, ny = a.ny, nz = a.nz;
  float bsum = 0.f;
  const int s_begin = blockIdx.x * a.segs_per_block;
  const int s_end = min(s_begin + a.segs_per_block, a.total_segs);
  for (int s = s_begin; s < s_end; ++s) {
    const int sx0 = (s % a.nseg_x) * kSeg;
    const int row = s / a.nseg_x;
    const int yy = row % ny, zz = row / ny;
    __syncthreads();
    // stage dY^T: dyT[co][p] = dY[pos(sx0 + p)][co]; 16-byte loads of 8
    // channels of one position (consecutive threads read consecutive
    // chunks), transposed on the way into LDS
    const long long rowpos = ((long long)zz * ny + yy) * nx;
    const int dch = a.dycp / 8;
    for (int q = threadIdx.x; q < kSeg * dch; q += 256) {
      const int p = q / dch, c8 = (q - p * dch) * 8;
      const int px = sx0 + p;
      u16 v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (px < nx) {
        const uint4 raw = *reinterpret_cast<const uint4*>(a.dy + (rowpos + px) * a.dycp + c8);
        __builtin_memcpy(v, &raw, 16);
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) dyT[(c8 + e) * kSeg + p] = v[e];
    }
    for (int q = threadIdx.x; q < (a.cout_pad - a.dycp) * kSeg; q += 256)
      dyT[a.dycp * kSeg + q] = 0;
    // stage X^T of the 9 neighbour rows, x from sx0-1 .. sx0+kSeg (each
    // row's positions are contiguous in memory: 16-byte loads of 8
    // channels, transposed into LDS)
    const int cch = cin / 8;
    const int per_row = xs * cch;
    for (int q = threadIdx.x; q < 9 * per_row; q += 256) {
      const int rr = q / per_row;
      const int rem = q - rr * per_row;
      const int p = rem / cch, c8 = (rem - p * cch) * 8;
      const int sz = zz + rr / 3 - 1, sy = yy + rr % 3 - 1, px = sx0 + p - 1;
      u16 v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      if (sz >= 0 && sz < nz && sy >= 0 && sy < ny && px >= 0 && px < nx) {
        const uint4 raw =
            *reinterpret_cast<const uint4*>(a.x + (((long long)sz * ny + sy) * nx + px) * cin + c8);
        __builtin_memcpy(v, &raw, 16);
      }
#pragma unroll
      for (int e = 0; e < 8; ++e) xT[(rr * cin + c8 + e) * xs + p] = v[e];
    }
    __syncthreads();
    // bias gradient: this block's sum of dY over its positions, per channel
    // (fixed order)
    if ((int)threadIdx.x < a.cout_pad) {
      const u16* row = dyT + threadIdx.x * kSeg;
      for (int p = 0; p < kSeg; ++p) bsum += bf2f(row[p]);
    }
    bf16x8 af[MT];
#pragma unroll
    for (int m = 0; m < MT; ++m) af[m] = load8(dyT + (m * 16 + r) * kSeg + 8 * g);
#pragma unroll
    for (int q = 0; q < kMaxTiles; ++q) {
      const int nt = wv + 4 * q;
      if (nt >= ntile) break;
      const int k = nt * 16 + r;  // this lane's B column
      const int tap = k / cin, c = k - tap * cin;
      bf16x8 bf;
      if (tap < 27) {
        const int rr = (tap / 9) * 3 + (tap / 3) % 3;
        const int dx = tap % 3 - 1;
        const u16* src = xT + (rr * cin + c) * xs + 8 * g + dx + 1;
        u16 tmp[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) tmp[v] = src[v];
        __builtin_memcpy(&bf, tmp, 16);
      } else {
        bf = bf16x8{0, 0, 0, 0, 0, 0, 0, 0};
      }
#pragma unroll
      for (int m = 0; m < MT; ++m)
        acc[m][q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bf, acc[m][q], 0, 0, 0);
    }
  }
  // partial sums: D[co = m*16 + 4g + v][k = nt*16 + r]
  float* out = a.part + (long long)blockIdx.x * a.cout_pad * a.kpad;
#pragma unroll
  for (int q = 0; q < kMaxTiles; ++q) {
    const int nt = wv + 4 * q;
    if (nt >= ntile) break;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
      for (int v = 0; v < 4; ++v)
        out[(long long)(m * 16 + 4 * g + v) * a.kpad + nt * 16 + r] = acc[m][q][v];
  }
  // bias partials after the gridDim.x weight partials
  if ((int)threadIdx.x < a.cout_pad)
    a.part[(long long)gridDim.x * a.cout_pad * a.kpad + (long long)blockIdx.x * a.cout_pad +
           threadIdx.x] = bsum;
}

// out[i] = sum_b part[b * n + i] for i < n_out, in a fixed order: four
// interleaved partial sums (so that four loads are in flight), then
// (s0 + s1) + (s2 + s3)
__global__ void wgrad_reduce_kernel(const float* part, int nparts, long long n, long long n_out,
                                    float* out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_out) return;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
  int b = 0;
  for (; b + 3 < nparts; b += 4) {
    s0 += part[(long long)b * n + i];
    s1 += part[(long long)(b + 1) * n + i];
    s2 += part[(long long)(b + 2) * n + i];
    s3 += part[(long long)(b + 3) * n + i];
  }
  for (; b < nparts; ++b) s0 += part[(long long)b * n + i];
  out[i] = (s0 + s1) + (s2 + s3);
}

__global__ void flip_weights_kernel(const u16* w, int cout, int cout_pad, int cin, int kpad,
                                    int co_t, int kpad_t, int rows_t, u16* wt) {
  // wt[ci][tap' * co_t + co] = w[co][(26 - tap') * cin + ci]
  const long long n = (long long)rows_t * kpad_t;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int ci = (int)(i / kpad_t), kk = (int)(i % kpad_t);
  const int tp = kk / co_t, co = kk % co_t;
  u16 v = 0;
  if (tp < 27 && co < cout && ci < cin) v = w[(long long)co * kpad + (26 - tp) * cin + ci];
  wt[i] = v;
}

// the cells of the box (bx0, by0, bz0) + [0, bnx) x [0, bny) x [0, bnz) of
// the grid, out in box order (the whole grid: the box (0, 0, 0, nx, ny, nz))
// one thread per box cell: x across the workgroup, blockIdx.y / .z the
// box row and plane (no 64-bit index division)
__global__ __launch_bounds__(128) void pack_fields_kernel(int nx, int ny, int bx0, int by0, int bz0,
                                                          int bnx, int bny, const float* f0,
                                                          const float* f1, const float* f2,
                                                          const float* f3, const float* f4,
                                                          const float* f5, const float* eps,
                                                          const float* scale, float eshift,
                                                          float escale, int flags, u16* out) {
  const int lx = blockIdx.x * 128 + threadIdx.x;
  if (lx >= bnx) return;
  const int ly = blockIdx.y, lz = blockIdx.z;
  const long long j = ((long long)lz * bny + ly) * bnx + lx;
  const long long i = ((long long)(bz0 + lz) * ny + (by0 + ly)) * nx + (bx0 + lx);
  const float* f[6] = {f0, f1, f2, f3, f4, f5};
  u16 v[8];
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    const float u = f[c][i] / scale[c];
    v[c] = f2bf((flags & CONV_PACK_ASINH) ? asinhf(u) : u);
  }
  v[6] = f2bf((eps[i] - eshift) / escale);
  v[7] = 0;
  uint4 pk;
  __builtin_memcpy(&pk, v, 16);
  *reinterpret_cast<uint4*>(out + j * 8) = pk;
}

__global__ void apply_correction_kernel(int nx, int ny, int nz, float* f0, float* f1, float* f2,
                                        float* f3, float* f4, float* f5, const float* y, int ys,
                                        const float* scale, float gain, int margin) {
  const long long n = (long long)nx * ny * nz;
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % nx), yy = (int)((i / nx) % ny), z = (int)(i / ((long long)nx * ny));
  if (x < margin || x >= nx - margin || yy < margin || yy >= ny - margin || z < margin ||
      z >= nz - margin)
    return;
  float* f[6] = {f0, f1, f2, f3, f4, f5};
#pragma unroll
  for (int c = 0; c < 6; ++c) f[c][i] += gain * scale[c] * y[i * ys + c];
}

// One closure correction in one pass (the same float operations, in the
// same order, as building y = net_gain (net - net(0)), y += far field,
// y[source boxes] = 0 and apply_correction): for the cells at least
// `margin` from every face,
//   v = net_gain * (yb - y0b) (the net's output and zero-state response on
//       the box (bx0, by0, bz0, bnx, bny, bnz), which holds those cells;
//       0 without yb) + s (the far-field operator's value; 0 without s),
//   with s dropped (mask_all == 0) or v = 0 (mask_all == 1) where mask != 0,
//   fields[c] += gain * scale[c] * v.
// (one thread per cell, x across the workgroup, blockIdx.y / .z the row and
// plane; VEC: yb, y0b and s with a channel stride of 8, read as two 16-byte
// loads per cell instead of six 4-byte ones)
template <bool VEC>
__global__ __launch_bounds__(128) void apply_closure_kernel(
    int nx, int ny, int nz, float* f0, float* f1, float* f2, float* f3, float* f4, float* f5,
    const float* s, int ss, const float* yb, const float* y0b, int ybs, int bx0, int by0, int bz0,
    int bnx, int bny, const unsigned char* mask, int mask_all, float ng, const float* scale,
    float gain, int margin) {
  const int x = blockIdx.x * 128 + threadIdx.x, yy = blockIdx.y, z = blockIdx.z;
  if (x < margin || x >= nx - margin || yy < margin || yy >= ny - margin || z < margin ||
      z >= nz - margin)
    return;
  const long long i = ((long long)z * ny + yy) * nx + x;
  const bool mk = mask && mask[i];
  const long long jb = yb ? ((long long)(z - bz0) * bny + (yy - by0)) * bnx + (x - bx0) : 0;
  float yv[8], y0v[8], sv[8];
  if (VEC) {
    if (yb) {
      const float4* p = reinterpret_cast<const float4*>(yb + jb * 8);
      const float4* q = reinterpret_cast<const float4*>(y0b + jb * 8);
      const float4 a = p[0], b = p[1], c = q[0], d = q[1];
      yv[0] = a.x, yv[1] = a.y, yv[2] = a.z, yv[3] = a.w, yv[4] = b.x, yv[5] = b.y;
      y0v[0] = c.x, y0v[1] = c.y, y0v[2] = c.z, y0v[3] = c.w, y0v[4] = d.x, y0v[5] = d.y;
    }
    if (s) {
      const float4* p = reinterpret_cast<const float4*>(s + i * 8);
      const float4 a = p[0], b = p[1];
      sv[0] = a.x, sv[1] = a.y, sv[2] = a.z, sv[3] = a.w, sv[4] = b.x, sv[5] = b.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      if (yb) yv[c] = yb[jb * ybs + c], y0v[c] = y0b[jb * ybs + c];
      if (s) sv[c] = s[i * ss + c];
    }
  }
  float* f[6] = {f0, f1, f2, f3, f4, f5};
#pragma unroll
  for (int c = 0; c < 6; ++c) {
    float v = yb ? (yv[c] - y0v[c]) * ng : 0.f;
    if (s) v += (mk && !mask_all) ? 0.f : sv[c];
    if (mk && mask_all) v = 0.f;
    f[c][i] += gain * scale[c] * v;
  }
}

// fields[c] += gain * corr[c] over the box (x0, y0, z0) + [0, bnx) x [0, bny)
// x [0, bnz); corr is [6][bnz][bny][bnx]
__global__ void patch_add_kernel(int nx, int ny, float* f0, float* f1, float* f2, float* f3,
                                 float* f4, float* f5, const float* corr, int x0, int y0, int z0,
                                 int bnx, int bny, int bnz, float gain) {
  const int nb = bnx * bny * bnz;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nb) return;
  const int lx = j % bnx, ly = (j / bnx) % bny, lz = j / (bnx * bny);
  const long long i = ((long long)(z0 + lz) * ny + (y0 + ly)) * nx + (x0 + lx);
  float* f[6] = {f0, f1, f2, f3, f4, f5};
#pragma unroll
  for (int c = 0; c < 6; ++c) f[c][i] = __builtin_fmaf(gain, corr[(long long)c * nb + j], f[c][i]);
}

// The closure's linear far-field operator, float32:
// y[i][co] += sum_{t, ci} (w[t][ci][0][co] + (eps[i] - 1) w[t][ci][1][co])
//             * f_ci(i + off(t)) / scale[ci]
// over the 27 taps off(t) = (dx, dy, dz) in {-1, 0, 1}^3, t = ((dz + 1) * 3
// + dy + 1) * 3 + dx + 1, zero outside the grid.  One thread per column pair
// (x, y, z0) and (x, y, z0 + 1): per field and (dx, dy) it loads the four
// planes z0 - 1 .. z0 + 2 once and feeds both cells, and each (tap, field)
// weight set (12 floats, contiguous, uniform across the workgroup: scalar
// loads) serves two cells.  Per field the taps are summed first (fused
// multiply-adds: one VALU instruction per term instead of a multiply and an
// add) and scaled once by 1 / scale[ci].
// (SET: y = the value instead of y += it; the cells of the box (bx0, by0,
// bz0) + [0, bnx) x [0, bny) x [0, bnz) only, y indexed on the whole grid)
template <bool SET>
__global__ __launch_bounds__(256) void stencil_f32_kernel(
    int nx, int ny, int nz, const float* __restrict__ f0, const float* __restrict__ f1,
    const float* __restrict__ f2, const float* __restrict__ f3, const float* __restrict__ f4,
    const float* __restrict__ f5, const float* __restrict__ eps,
    const float* __restrict__ scale, const float* __restrict__ w, float* __restrict__ y,
    int ys, int bx0, int by0, int bz0, int bnx, int bny, int bnz) {
  const int nzp = (bnz + 1) / 2;
  const long long npairs = (long long)bnx * bny * nzp;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npairs) return;
  const int x = bx0 + (int)(p % bnx), yy = by0 + (int)((p / bnx) % bny);
  const int z0 = bz0 + 2 * (int)(p / ((long long)bnx * bny));
  const bool two = z0 + 1 < bz0 + bnz;
  const long long plane = (long long)nx * ny;
  const long long i0 = (long long)z0 * plane + (long long)yy * nx + x;
  const float* f[6] = {f0, f1, f2, f3, f4, f5};
  float a0[6], b0[6], a1[6], b1[6];
#pragma unroll
  for (int co = 0; co < 6; ++co) a0[co] = b0[co] = a1[co] = b1[co] = 0.f;
  // the field loop stays rolled: 81 instead of 168 VGPRs, 5 wavefronts per
  // SIMD instead of 3 (0.211 against 0.252 ms at 128^3)
#pragma unroll 1
  for (int ci = 0; ci < 6; ++ci) {
    float pa0[6], pb0[6], pa1[6], pb1[6];
#pragma unroll
    for (int co = 0; co < 6; ++co) pa0[co] = pb0[co] = pa1[co] = pb1[co] = 0.f;
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int dx = q % 3 - 1, dy = q / 3 - 1;
      const int sx = x + dx, sy = yy + dy;
      if (sx < 0 || sx >= nx || sy < 0 || sy >= ny) continue;
      // planes z0 - 1 .. z0 + 2 of column (sx, sy); zero outside the grid
      float v[4];
#pragma unroll
      for (int pz = 0; pz < 4; ++pz) {
        const int sz = z0 - 1 + pz;
        v[pz] = (sz >= 0 && sz < nz) ? f[ci][i0 + (long long)(pz - 1) * plane + dy * nx + dx]
                                     : 0.f;
      }
#pragma unroll
      for (int dz = 0; dz < 3; ++dz) {
        const float* wt = w + ((dz * 9 + q) * 6 + ci) * 12;
#pragma unroll
        for (int co = 0; co < 6; ++co) {
          pa0[co] = __builtin_fmaf(wt[co], v[dz], pa0[co]);
          pb0[co] = __builtin_fmaf(wt[6 + co], v[dz], pb0[co]);
          pa1[co] = __builtin_fmaf(wt[co], v[dz + 1], pa1[co]);
          pb1[co] = __builtin_fmaf(wt[6 + co], v[dz + 1], pb1[co]);
        }
      }
    }
    const float inv = 1.0f / scale[ci];
#pragma unroll
    for (int co = 0; co < 6; ++co) {
      a0[co] += pa0[co] * inv;
      b0[co] += pb0[co] * inv;
      a1[co] += pa1[co] * inv;
      b1[co] += pb1[co] * inv;
    }
  }
  const float e0 = eps[i0] - 1.0f;
#pragma unroll
  for (int co = 0; co < 6; ++co) {
    const float v = a0[co] + e0 * b0[co];
    if (SET) y[i0 * ys + co] = v;
    else y[i0 * ys + co] += v;
  }
  if (two) {
    const long long i1 = i0 + plane;
    const float e1 = eps[i1] - 1.0f;
#pragma unroll
    for (int co = 0; co < 6; ++co) {
      const float v = a1[co] + e1 * b1[co];
      if (SET) y[i1 * ys + co] = v;
      else y[i1 * ys + co] += v;
    }
  }
}

// The same operator, four cells (x, y, z0 .. z0 + 3) per thread: the
// workgroup first stages the weights divided by scale[ci] in LDS (read back
// as wave-uniform 16-byte broadcasts), so the taps of all six fields sum
// straight into the outputs (no per-field partial sums), and each loaded
// input value feeds up to three of the four cells.  The float32 sums are
// those of the two-cell kernel up to rounding (w / scale[ci] in place of
// (sum w f) / scale[ci]).
template <bool SET>
__global__ __launch_bounds__(256) void stencil4_f32_kernel(
    int nx, int ny, int nz, const float* __restrict__ f0, const float* __restrict__ f1,
    const float* __restrict__ f2, const float* __restrict__ f3, const float* __restrict__ f4,
    const float* __restrict__ f5, const float* __restrict__ eps,
    const float* __restrict__ scale, const float* __restrict__ w, float* __restrict__ y,
    int ys, int bx0, int by0, int bz0, int bnx, int bny, int bnz) {
  __shared__ __attribute__((aligned(16))) float ws[27 * 6 * 12];
  for (int e = threadIdx.x; e < 27 * 6 * 12; e += 256) ws[e] = w[e] * (1.0f / scale[(e / 12) % 6]);
  __syncthreads();
  const int nzq = (bnz + 3) / 4;
  const long long nq = (long long)bnx * bny * nzq;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nq) return;
  const int lx = (int)(p % bnx);
  const int rest = (int)(p / bnx);
  const int x = bx0 + lx, yy = by0 + rest % bny;
  const int z0 = bz0 + 4 * (rest / bny);
  const int ncell = min(4, bz0 + bnz - z0);
  const long long plane = (long long)nx * ny;
  const long long i0 = (long long)z0 * plane + (long long)yy * nx + x;
  const float* f[6] = {f0, f1, f2, f3, f4, f5};
  float a[4][6], b[4][6];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int co = 0; co < 6; ++co) a[c][co] = b[c][co] = 0.f;
#pragma unroll 1
  for (int ci = 0; ci < 6; ++ci) {
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int dx = q % 3 - 1, dy = q / 3 - 1;
      const int sx = x + dx, sy = yy + dy;
      if (sx < 0 || sx >= nx || sy < 0 || sy >= ny) continue;
      // planes z0 - 1 .. z0 + 4 of column (sx, sy); zero outside the grid
      float v[6];
#pragma unroll
      for (int pz = 0; pz < 6; ++pz) {
        const int sz = z0 - 1 + pz;
        v[pz] = (sz >= 0 && sz < nz) ? f[ci][i0 + (long long)(pz - 1) * plane + dy * nx + dx]
                                     : 0.f;
      }
#pragma unroll
      for (int dz = 0; dz < 3; ++dz) {
        const float4* wt = reinterpret_cast<const float4*>(ws + ((dz * 9 + q) * 6 + ci) * 12);
        const float4 w0 = wt[0], w1 = wt[1], w2 = wt[2];
        const float wa[6] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y};
        const float wb[6] = {w1.z, w1.w, w2.x, w2.y, w2.z, w2.w};
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int co = 0; co < 6; ++co) {
            a[c][co] = __builtin_fmaf(wa[co], v[c + dz], a[c][co]);
            b[c][co] = __builtin_fmaf(wb[co], v[c + dz], b[c][co]);
          }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c >= ncell) break;
    const long long i = i0 + c * plane;
    const float e = eps[i] - 1.0f;
#pragma unroll
    for (int co = 0; co < 6; ++co) {
      const float v = a[c][co] + e * b[c][co];
      if (SET) y[i * ys + co] = v;
      else y[i * ys + co] += v;
    }
  }
}

// ---------------------------------------------------------------------------
// Near-source patch of the closure: restriction of a 2x-refined local grid
// to the coarse box around a source, and the coarse box's correction.  The
// same float32 operations in the same order as the torch restriction in
// jaxphys_amd/closure.py (restrict_e / restrict_h, the mean of two H
// restrictions), so the two give the same bits.  One thread per coarse cell
// of the side^3 box at coarse (x0, y0, z0); fine index of coarse (K, J, I)
// is (2K, 2J, 2I).
// ---------------------------------------------------------------------------

__device__ __forceinline__ long long fidx(int nxf, int nyf, int z, int y, int x) {
  return ((long long)z * nyf + y) * nxf + x;
}

// coarse Hx, Hy, Hz of the box (each the mean of four fine values)
__device__ __forceinline__ void restrict_h_cell(const float* __restrict__ hx,
                                                const float* __restrict__ hy,
                                                const float* __restrict__ hz, int nxf, int nyf,
                                                int K, int J, int I, float& ox, float& oy,
                                                float& oz) {
  const int z = 2 * K, y = 2 * J, x = 2 * I;
  ox = 0.25f * (((hx[fidx(nxf, nyf, z, y, x)] + hx[fidx(nxf, nyf, z, y + 1, x)]) +
                 hx[fidx(nxf, nyf, z + 1, y, x)]) + hx[fidx(nxf, nyf, z + 1, y + 1, x)]);
  oy = 0.25f * (((hy[fidx(nxf, nyf, z, y, x)] + hy[fidx(nxf, nyf, z, y, x + 1)]) +
                 hy[fidx(nxf, nyf, z + 1, y, x)]) + hy[fidx(nxf, nyf, z + 1, y, x + 1)]);
  oz = 0.25f * (((hz[fidx(nxf, nyf, z, y, x)] + hz[fidx(nxf, nyf, z, y, x + 1)]) +
                 hz[fidx(nxf, nyf, z, y + 1, x)]) + hz[fidx(nxf, nyf, z, y + 1, x + 1)]);
}

__global__ __launch_bounds__(256) void patch_restrict_h_kernel(
    const float* __restrict__ hx, const float* __restrict__ hy, const float* __restrict__ hz,
    int nxf, int nyf, int x0, int y0, int z0, int side, float* __restrict__ out) {
  const int n = side * side * side;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = t % side, j = (t / side) % side, k = t / (side * side);
  float a, b, c;
  restrict_h_cell(hx, hy, hz, nxf, nyf, z0 + k, y0 + j, x0 + i, a, b, c);
  out[t] = a;
  out[n + t] = b;
  out[2 * n + t] = c;
}

__global__ __launch_bounds__(256) void patch_correct_kernel(
    const float* __restrict__ ex, const float* __restrict__ ey, const float* __restrict__ ez,
    const float* __restrict__ hx, const float* __restrict__ hy, const float* __restrict__ hz,
    const float* __restrict__ hprev, int nxf, int nyf, float* __restrict__ c0,
    float* __restrict__ c1, float* __restrict__ c2, float* __restrict__ c3,
    float* __restrict__ c4, float* __restrict__ c5, int nxc, int nyc, int x0, int y0, int z0,
    int side, float* __restrict__ corr) {
  const int n = side * side * side;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = t % side, j = (t / side) % side, k = t / (side * side);
  const int K = z0 + k, J = y0 + j, I = x0 + i;
  const int z = 2 * K, y = 2 * J, x = 2 * I;
  float st[6];
  st[0] = 0.5f * (ex[fidx(nxf, nyf, z, y, x)] + ex[fidx(nxf, nyf, z, y, x + 1)]);
  st[1] = 0.5f * (ey[fidx(nxf, nyf, z, y, x)] + ey[fidx(nxf, nyf, z, y + 1, x)]);
  st[2] = 0.5f * (ez[fidx(nxf, nyf, z, y, x)] + ez[fidx(nxf, nyf, z + 1, y, x)]);
  float h0, h1, h2;
  restrict_h_cell(hx, hy, hz, nxf, nyf, K, J, I, h0, h1, h2);
  st[3] = 0.5f * (hprev[t] + h0);
  st[4] = 0.5f * (hprev[n + t] + h1);
  st[5] = 0.5f * (hprev[2 * n + t] + h2);
  float* cf[6] = {c0, c1, c2, c3, c4, c5};
  const long long o = ((long long)K * nyc + J) * nxc + I;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const float x_ = cf[q][o];
    const float d = st[q] - x_;
    cf[q][o] = x_ + d;
    corr[q * n + t] = d;
  }
}

int kpad_of(int cin) { return ((27 * cin + 31) / 32) * 32; }

// JAXPHYS_AMD_CONV_NOSHIFT=1 turns the shifted-fragment loop off (same
// results; for comparison).
bool shift_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("JAXPHYS_AMD_CONV_NOSHIFT");
    return !(e && e[0] == '1');
  }();
  return on;
}

// JAXPHYS_AMD_CONV_ROWS=0 turns the row kernel off (the general kernel then
// runs every layer; for comparison).
bool rows_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("JAXPHYS_AMD_CONV_ROWS");
    return !(e && e[0] == '0');
  }();
  return on;
}

// JAXPHYS_AMD_CONV_SLAB=0 turns the slab kernel off (32-channel layers then
// take the row kernel).
bool slab_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("JAXPHYS_AMD_CONV_SLAB");
    return !(e && e[0] == '0');
  }();
  return on;
}

// JAXPHYS_AMD_STENCIL4=0 takes the two-cell far-field stencil kernel
// instead of the four-cell one (for comparison).
bool stencil4_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("JAXPHYS_AMD_STENCIL4");
    return !(e && e[0] == '0');
  }();
  return on;
}

// JAXPHYS_AMD_CONV_PAIRX=0 turns the paired-tap slab kernel off (float32
// layers of at most 8 channels then pad them to 16; for comparison).
bool pairx_enabled() {
  static const bool on = [] {
    const char* e = std::getenv("JAXPHYS_AMD_CONV_PAIRX");
    return !(e && e[0] == '0');
  }();
  return on;
}

template <int EPI, int MT>
int slab_launch(const ConvArgs& a, dim3 grid, hipStream_t st) {
  if constexpr (EPI == CONV_EPI_LINEAR_F32 && MT == 1) {
    if (a.cout <= 8 && a.ycp <= 8 && pairx_enabled()) {
      hipLaunchKernelGGL((conv_slab_kernel<EPI, MT, true>), grid, dim3(1024), 0, st, a);
      return 0;
    }
  }
  hipLaunchKernelGGL((conv_slab_kernel<EPI, MT>), grid, dim3(1024), 0, st, a);
  return 0;
}

size_t row_shm(int cin, int mt, int epi) {
  const size_t kp = cin == 32 ? 27 * 32 : 9 * 32;
  return (size_t)mt * 16 * (kp + 8) * sizeof(u16) +
         (epi == CONV_EPI_GELU ? 2 * (size_t)kRowLutS * sizeof(u16) : 0);
}

}  // namespace conv

using namespace conv;

extern "C" {

const char* conv3d_last_error(void) { return g_err.c_str(); }

int conv3d_fwd(const ConvDesc* d, const void* x, const void* w, const float* bias, int32_t epi,
               void* z_out, const void* aux, void* y, void* stream) {
  if (!d || !x || !w || !y) return fail(-1, "null argument");
  if (d->cin_pad <= 0 || d->cin_pad % 8 || d->cout_pad % 16 || d->cout_pad <= 0 ||
      d->cout_pad > 32 || d->cout > d->cout_pad || d->y_cpad < d->cout || d->y_cpad > d->cout_pad)
    return fail(-1, "unsupported channel configuration");
  if (d->nx < 1 || d->ny < 1 || d->nz < 1) return fail(-1, "bad grid");
  if (epi == CONV_EPI_GELU_BWD && !aux) return fail(-1, "GELU backward needs the pre-activation");
  if ((epi == CONV_EPI_GELU || epi == CONV_EPI_GELU_BWD || epi == CONV_EPI_LINEAR_BF16) &&
      d->y_cpad % 8)
    return fail(-1, "bf16 outputs need y_cpad multiple of 8");
  ConvArgs a;
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.cin = d->cin_pad;
  a.cout = d->cout;
  a.cout_pad = d->cout_pad;
  a.ycp = d->y_cpad;
  a.kpad = kpad_of(d->cin_pad);
  a.x = (const u16*)x;
  a.w = (const u16*)w;
  a.bias = bias;
  a.zout = (u16*)z_out;
  a.aux = (const u16*)aux;
  a.y = y;
  constexpr int NT = 4;
  constexpr int WAVES = 16;
  const int gx = (d->nx + 16 * NT - 1) / (16 * NT);
  const int gy = (d->ny + WAVES - 1) / WAVES;
  long long blocks_xy = (long long)gx * gy;
  // ~512 workgroups: one 1024-thread workgroup fits a CU at a time (its
  // registers), and each stages the layer's weights before computing, so
  // more planes per workgroup amortise the staging (128^3, 32 -> 32 -> 32
  // -> 6 net: 1 plane 0.596 ms, 2 planes 0.572, 4 planes 0.568, 8 planes
  // 0.569 per inference)
  int zchunk = (int)((blocks_xy * d->nz + 511) / 512);
  zchunk = zchunk < 1 ? 1 : (zchunk > 8 ? 8 : zchunk);
  a.zchunk = zchunk;
  const dim3 grid(gx, gy, (d->nz + zchunk - 1) / zchunk);
  const size_t shm = (size_t)d->cout_pad * (a.kpad + 8) * sizeof(u16);
  if (shm > 160 * 1024) return fail(-1, "weights exceed LDS");
  hipStream_t st = (hipStream_t)stream;
  const bool shift = d->cin_pad % 32 == 0 && shift_enabled();
  auto al8 = [](const void* p) { return ((uintptr_t)p & 7) == 0; };
  const bool rows = rows_enabled() && (d->cin_pad == 32 || d->cin_pad == 8) &&
                    (epi == CONV_EPI_LINEAR_F32 ||
                     (d->y_cpad % 4 == 0 && al8(y) && al8(z_out) && al8(aux)));
  // slab kernel: 64 x 8 x (2 z) per block step, about one block per CU
  const bool slab = rows && d->cin_pad == 32 && slab_enabled() &&
                    (long long)d->nx * d->ny * 64 < (1LL << 31);
  const int sgx = (d->nx + 63) / 64, sgy = (d->ny + 7) / 8;
  int szc = (int)(((long long)sgx * sgy * d->nz + 255) / 256);
  szc = szc < 2 ? 2 : szc + (szc & 1);
  ConvArgs as = a;
  as.zchunk = szc;
  const dim3 sgrid(sgx, sgy, (d->nz + szc - 1) / szc);
  // rolling-slab kernel for 8 channels: 64 x 16 rows per block, about one
  // block per CU
  const bool roll = rows && d->cin_pad == 8 && slab_enabled() &&
                    (long long)d->nx * d->ny * 16 < (1LL << 31);
  const int rgx = (d->nx + 63) / 64, rgy = (d->ny + 15) / 16;
  int rzc = (int)(((long long)rgx * rgy * d->nz + 255) / 256);
  rzc = rzc < 1 ? 1 : rzc;
  ConvArgs ar = a;
  ar.zchunk = rzc;
  const dim3 rgrid(rgx, rgy, (d->nz + rzc - 1) / rzc);
#define CONV_LAUNCH(E, MT)                                                                   \
  do {                                                                                       \
    if (slab) {                                                                              \
      const int rc_ = slab_launch<E, MT>(as, sgrid, st);                                     \
      if (rc_) return rc_;                                                                   \
    } else if (roll) {                                                                       \
      hipLaunchKernelGGL((conv_roll8_kernel<E, MT>), rgrid, dim3(1024), 0, st, ar);          \
    } else if (rows && d->cin_pad == 32)                                                     \
      hipLaunchKernelGGL((conv_row_kernel<E, MT, 32>), grid, dim3(1024),                     \
                         row_shm(32, MT, E), st, a);                                         \
    else if (rows)                                                                           \
      hipLaunchKernelGGL((conv_row_kernel<E, MT, 8>), grid, dim3(1024),                      \
                         row_shm(8, MT, E), st, a);                                          \
    else if (shift)                                                                          \
      hipLaunchKernelGGL((conv_fwd_kernel<E, NT, MT, WAVES, true>), grid, dim3(64 * WAVES),  \
                         shm, st, a);                                                        \
    else if (d->cin_pad == 8)                                                                \
      hipLaunchKernelGGL((conv_fwd_kernel<E, NT, MT, WAVES, false, 8>), grid,                \
                         dim3(64 * WAVES), shm, st, a);                                      \
    else                                                                                     \
      hipLaunchKernelGGL((conv_fwd_kernel<E, NT, MT, WAVES, false>), grid, dim3(64 * WAVES), \
                         shm, st, a);                                                        \
  } while (0)
  const int mt = d->cout_pad / 16;
  switch (epi) {
    case CONV_EPI_GELU:
      if (mt == 2) CONV_LAUNCH(CONV_EPI_GELU, 2); else CONV_LAUNCH(CONV_EPI_GELU, 1);
      break;
    case CONV_EPI_LINEAR_F32:
      if (mt == 2) CONV_LAUNCH(CONV_EPI_LINEAR_F32, 2); else CONV_LAUNCH(CONV_EPI_LINEAR_F32, 1);
      break;
    case CONV_EPI_GELU_BWD:
      if (mt == 2) CONV_LAUNCH(CONV_EPI_GELU_BWD, 2); else CONV_LAUNCH(CONV_EPI_GELU_BWD, 1);
      break;
    case CONV_EPI_LINEAR_BF16:
      if (mt == 2) CONV_LAUNCH(CONV_EPI_LINEAR_BF16, 2); else CONV_LAUNCH(CONV_EPI_LINEAR_BF16, 1);
      break;
    default: return fail(-1, "unknown epilogue");
  }
#undef CONV_LAUNCH
  CONV_CHECK(hipGetLastError());
  return 0;
}

static void wgrad_shape(const ConvDesc* d, int& nseg_x, int& total, int& spb, int& nblocks) {
  nseg_x = (d->nx + kSeg - 1) / kSeg;
  total = nseg_x * d->ny * d->nz;
  nblocks = total < 1024 ? total : 1024;
  spb = (total + nblocks - 1) / nblocks;
  nblocks = (total + spb - 1) / spb;
}

int64_t conv3d_wgrad_work_floats(const ConvDesc* d) {
  if (!d) return 0;
  int nseg_x, total, spb, nb;
  wgrad_shape(d, nseg_x, total, spb, nb);
  const long long kp = kpad_of(d->cin_pad);
  return (long long)nb * d->cout_pad * kp + (long long)nb * d->cout_pad;
}

int conv3d_wgrad(const ConvDesc* d, const void* x, const void* dy, int32_t dy_cpad, float* work,
                 float* dw, float* db, void* stream) {
  if (!d || !x || !dy || !work || !dw) return fail(-1, "null argument");
  if (d->cin_pad % 8 || d->cout_pad % 16 || d->cout_pad > 32 || dy_cpad > d->cout_pad ||
      dy_cpad < d->cout || dy_cpad % 8)
    return fail(-1, "unsupported channel configuration");
  if (((uintptr_t)x | (uintptr_t)dy) % 16)
    return fail(-1, "x and dy must be 16-byte aligned");
  const int kpad = kpad_of(d->cin_pad);
  if (kpad / 16 > 4 * 16) return fail(-1, "too many input channels for the weight gradient");
  int nseg_x, total, spb, nb;
  wgrad_shape(d, nseg_x, total, spb, nb);
  WgradArgs a;
  a.nx = d->nx;
  a.ny = d->ny;
  a.nz = d->nz;
  a.cin = d->cin_pad;
  a.cout_pad = d->cout_pad;
  a.dycp = dy_cpad;
  a.kpad = kpad;
  a.nseg_x = nseg_x;
  a.segs_per_block = spb;
  a.total_segs = total;
  a.x = (const u16*)x;
  a.dy = (const u16*)dy;
  a.part = work;
  hipStream_t st = (hipStream_t)stream;
  const size_t shm = ((size_t)d->cout_pad * kSeg + 9 * (size_t)d->cin_pad * (kSeg + 2)) * sizeof(u16);
  if (d->cout_pad == 32)
    hipLaunchKernelGGL((conv_wgrad_kernel<2>), dim3(nb), dim3(256), shm, st, a);
  else
    hipLaunchKernelGGL((conv_wgrad_kernel<1>), dim3(nb), dim3(256), shm, st, a);
  CONV_CHECK(hipGetLastError());
  const long long n = (long long)d->cout_pad * kpad;
  hipLaunchKernelGGL(wgrad_reduce_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                     (const float*)work, nb, n, n, dw);
  CONV_CHECK(hipGetLastError());
  if (db) {
    // the kernel's per-block bias partials follow its weight partials
    hipLaunchKernelGGL(wgrad_reduce_kernel, dim3(1), dim3(64), 0, st,
                       (const float*)(work + (long long)nb * n), nb, (long long)d->cout_pad,
                       (long long)d->cout, db);
    CONV_CHECK(hipGetLastError());
  }
  return 0;
}

int conv3d_flip_weights(const ConvDesc* d, const void* w, int32_t dy_cpad, void* wt, void* stream) {
  if (!d || !w || !wt) return fail(-1, "null argument");
  const int kpad = kpad_of(d->cin_pad);
  const int co_t = ((dy_cpad + 7) / 8) * 8;
  const int kpad_t = kpad_of(co_t);
  const int rows_t = ((d->cin_pad + 15) / 16) * 16;
  const long long n = (long long)rows_t * kpad_t;
  hipLaunchKernelGGL(flip_weights_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const u16*)w, d->cout, d->cout_pad, d->cin_pad, kpad,
                     co_t, kpad_t, rows_t, (u16*)wt);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_pack_fields(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* eps,
                       const float* scale, float eps_shift, float eps_scale, int32_t flags,
                       void* out, void* stream) {
  return conv3d_pack_fields_box(nx, ny, nz, fields, eps, scale, eps_shift, eps_scale, flags, 0, 0,
                                0, nx, ny, nz, out, stream);
}

int conv3d_pack_fields_box(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           const float* eps, const float* scale, float eps_shift, float eps_scale,
                           int32_t flags, int32_t x0, int32_t y0, int32_t z0, int32_t bnx,
                           int32_t bny, int32_t bnz, void* out, void* stream) {
  if (!fields || !eps || !scale || !out) return fail(-1, "null argument");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(-1, "null field");
  if (bnx <= 0 || bny <= 0 || bnz <= 0 || x0 < 0 || y0 < 0 || z0 < 0 || x0 + bnx > nx ||
      y0 + bny > ny || z0 + bnz > nz)
    return fail(-1, "box outside the grid");
  if (bny > 65535 || bnz > 65535) return fail(-1, "box too large");
  hipLaunchKernelGGL(pack_fields_kernel, dim3((unsigned)((bnx + 127) / 128), bny, bnz), dim3(128),
                     0, (hipStream_t)stream, nx, ny, x0, y0, z0, bnx, bny, fields[0], fields[1],
                     fields[2], fields[3], fields[4], fields[5], eps, scale, eps_shift, eps_scale,
                     flags, (u16*)out);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_apply_closure(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* s,
                         int32_t sstride, const float* yb, const float* y0b, int32_t ybstride,
                         const int32_t* box, const unsigned char* mask, int32_t mask_all,
                         float net_gain, const float* scale, float gain, int32_t margin,
                         void* stream) {
  if (!fields || !scale || margin < 0) return fail(-1, "bad argument");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(-1, "null field");
  if (s && sstride < 6) return fail(-1, "far-field stride below 6");
  int b[6] = {0, 0, 0, 0, 0, 0};
  if (yb) {
    if (!y0b || !box || ybstride < 6) return fail(-1, "net output without its box");
    for (int k = 0; k < 6; ++k) b[k] = box[k];
    // the box must hold every cell `margin` or more from the faces
    const int lo[3] = {margin, margin, margin}, hi[3] = {nx - margin, ny - margin, nz - margin};
    for (int k = 0; k < 3; ++k)
      if (hi[k] > lo[k] && (b[k] > lo[k] || b[k] + b[k + 3] < hi[k]))
        return fail(-1, "net box does not cover the corrected cells");
  }
  if (ny > 65535 || nz > 65535) return fail(-1, "grid too large");
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool vec = (!yb || (ybstride == 8 && al16(yb) && al16(y0b))) &&
                   (!s || (sstride == 8 && al16(s)));
  const dim3 grid((unsigned)((nx + 127) / 128), ny, nz);
  if (vec)
    hipLaunchKernelGGL(apply_closure_kernel<true>, grid, dim3(128), 0, (hipStream_t)stream, nx, ny,
                       nz, fields[0], fields[1], fields[2], fields[3], fields[4], fields[5], s,
                       sstride, yb, y0b, ybstride, b[0], b[1], b[2], b[3], b[4], mask, mask_all,
                       net_gain, scale, gain, margin);
  else
    hipLaunchKernelGGL(apply_closure_kernel<false>, grid, dim3(128), 0, (hipStream_t)stream, nx,
                       ny, nz, fields[0], fields[1], fields[2], fields[3], fields[4], fields[5], s,
                       sstride, yb, y0b, ybstride, b[0], b[1], b[2], b[3], b[4], mask, mask_all,
                       net_gain, scale, gain, margin);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_patch_add(int32_t nx, int32_t ny, int32_t nz, float* const* fields, const float* corr,
                     int32_t x0, int32_t y0, int32_t z0, int32_t bnx, int32_t bny, int32_t bnz,
                     float gain, void* stream) {
  if (!fields || !corr) return fail(-1, "bad argument");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(-1, "null field");
  if (bnx <= 0 || bny <= 0 || bnz <= 0 || x0 < 0 || y0 < 0 || z0 < 0 || x0 + bnx > nx ||
      y0 + bny > ny || z0 + bnz > nz)
    return fail(-1, "box outside the grid");
  const int nb = bnx * bny * bnz;
  hipLaunchKernelGGL(patch_add_kernel, dim3((unsigned)((nb + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, nx, ny, fields[0], fields[1], fields[2], fields[3],
                     fields[4], fields[5], corr, x0, y0, z0, bnx, bny, bnz, gain);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_apply_correction(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                            const float* y, int32_t ystride, const float* scale, float gain,
                            int32_t margin, void* stream) {
  if (!fields || !y || !scale || ystride < 6 || margin < 0) return fail(-1, "bad argument");
  const long long n = (long long)nx * ny * nz;
  hipLaunchKernelGGL(apply_correction_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, nx, ny, nz, fields[0], fields[1], fields[2], fields[3],
                     fields[4], fields[5], y, ystride, scale, gain, margin);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_stencil_f32(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                       const float* eps, const float* scale, const float* w, float* y,
                       int32_t ystride, void* stream) {
  if (!fields || !eps || !scale || !w || !y || ystride < 6 || nx <= 0 || ny <= 0 || nz <= 0)
    return fail(-1, "bad argument");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(-1, "null field");
  if (stencil4_enabled()) {
    const long long nq = (long long)nx * ny * ((nz + 3) / 4);
    hipLaunchKernelGGL(stencil4_f32_kernel<false>, dim3((unsigned)((nq + 255) / 256)), dim3(256),
                       0, (hipStream_t)stream, nx, ny, nz, fields[0], fields[1], fields[2],
                       fields[3], fields[4], fields[5], eps, scale, w, y, ystride, 0, 0, 0, nx, ny,
                       nz);
  } else {
    const long long npairs = (long long)nx * ny * ((nz + 1) / 2);
    hipLaunchKernelGGL(stencil_f32_kernel<false>, dim3((unsigned)((npairs + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, nx, ny, nz, fields[0], fields[1],
                       fields[2], fields[3], fields[4], fields[5], eps, scale, w, y, ystride, 0, 0,
                       0, nx, ny, nz);
  }
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_stencil_f32_set(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           const float* eps, const float* scale, const float* w, float* y,
                           int32_t ystride, const int32_t* box, void* stream) {
  if (!fields || !eps || !scale || !w || !y || ystride < 6 || nx <= 0 || ny <= 0 || nz <= 0)
    return fail(-1, "bad argument");
  for (int c = 0; c < 6; ++c)
    if (!fields[c]) return fail(-1, "null field");
  int b[6] = {0, 0, 0, nx, ny, nz};
  if (box) {
    for (int k = 0; k < 6; ++k) b[k] = box[k];
    if (b[3] <= 0 || b[4] <= 0 || b[5] <= 0 || b[0] < 0 || b[1] < 0 || b[2] < 0 ||
        b[0] + b[3] > nx || b[1] + b[4] > ny || b[2] + b[5] > nz)
      return fail(-1, "box outside the grid");
  }
  if (stencil4_enabled()) {
    const long long nq = (long long)b[3] * b[4] * ((b[5] + 3) / 4);
    hipLaunchKernelGGL(stencil4_f32_kernel<true>, dim3((unsigned)((nq + 255) / 256)), dim3(256),
                       0, (hipStream_t)stream, nx, ny, nz, fields[0], fields[1], fields[2],
                       fields[3], fields[4], fields[5], eps, scale, w, y, ystride, b[0], b[1], b[2],
                       b[3], b[4], b[5]);
  } else {
    const long long npairs = (long long)b[3] * b[4] * ((b[5] + 1) / 2);
    hipLaunchKernelGGL(stencil_f32_kernel<true>, dim3((unsigned)((npairs + 255) / 256)),
                       dim3(256), 0, (hipStream_t)stream, nx, ny, nz, fields[0], fields[1],
                       fields[2], fields[3], fields[4], fields[5], eps, scale, w, y, ystride, b[0],
                       b[1], b[2], b[3], b[4], b[5]);
  }
  CONV_CHECK(hipGetLastError());
  return 0;
}

static bool box_inside(int x0, int y0, int z0, int side, int nx, int ny, int nz) {
  return side > 0 && x0 >= 0 && y0 >= 0 && z0 >= 0 && x0 + side <= nx && y0 + side <= ny &&
         z0 + side <= nz;
}

int conv3d_patch_restrict_h(const float* const* hf, int32_t nxf, int32_t nyf, int32_t nzf,
                            int32_t x0, int32_t y0, int32_t z0, int32_t side, float* out,
                            void* stream) {
  if (!hf || !hf[0] || !hf[1] || !hf[2] || !out) return fail(-1, "bad argument");
  if (nxf % 2 || nyf % 2 || nzf % 2) return fail(-1, "fine grid sides must be even");
  if (!box_inside(x0, y0, z0, side, nxf / 2, nyf / 2, nzf / 2))
    return fail(-1, "box outside the coarse grid of the fine one");
  const int n = side * side * side;
  hipLaunchKernelGGL(patch_restrict_h_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, hf[0], hf[1], hf[2], nxf, nyf, x0, y0, z0, side, out);
  CONV_CHECK(hipGetLastError());
  return 0;
}

int conv3d_patch_correct(const float* const* ef, const float* const* hf, const float* hprev,
                         int32_t nxf, int32_t nyf, int32_t nzf, float* const* coarse,
                         int32_t nxc, int32_t nyc, int32_t nzc, int32_t x0, int32_t y0,
                         int32_t z0, int32_t side, float* corr, void* stream) {
  if (!ef || !hf || !hprev || !coarse || !corr) return fail(-1, "bad argument");
  for (int c = 0; c < 3; ++c)
    if (!ef[c] || !hf[c]) return fail(-1, "null fine field");
  for (int c = 0; c < 6; ++c)
    if (!coarse[c]) return fail(-1, "null coarse field");
  if (nxf % 2 || nyf % 2 || nzf % 2) return fail(-1, "fine grid sides must be even");
  if (!box_inside(x0, y0, z0, side, nxf / 2, nyf / 2, nzf / 2) ||
      !box_inside(x0, y0, z0, side, nxc, nyc, nzc))
    return fail(-1, "box outside the coarse grid");
  const int n = side * side * side;
  hipLaunchKernelGGL(patch_correct_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, ef[0], ef[1], ef[2], hf[0], hf[1], hf[2], hprev, nxf,
                     nyf, coarse[0], coarse[1], coarse[2], coarse[3], coarse[4], coarse[5], nxc,
                     nyc, x0, y0, z0, side, corr);
  CONV_CHECK(hipGetLastError());
  return 0;
}

}  // extern "C"
