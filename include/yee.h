/*
 * jaxphys_amd -- C ABI of the MI355X (gfx950) 3D Yee FDTD library
 * (libyee_hip.so).
 *
 * The reference (sushaan-k/jaxphys, package neurosim) has no FFI: its hot
 * path is the Python closure `step` inside `EMGrid3D.simulate`
 * (src/neurosim/em/fdtd3d.py:280-379) driven by `jax.lax.scan`
 * (fdtd3d.py:393).  This header is the seam that replaces that scan: the
 * Python `EMGrid3D` (same construction/run API as fdtd3d.py:71-417) builds
 * the coefficient arrays once and calls these entry points through ctypes.
 *
 * Storage: every per-cell array is float32, x fastest,
 *     idx(i, j, k) = (k * ny + j) * nx + i,
 * where (i, j, k) are the reference's (x, y, z) indices F[i, j, k]
 * (fdtd3d.py:212-217, 240).  All pointers are device pointers owned by the
 * caller (torch tensors); the library borrows them for the duration of a
 * call.  All launches go on the caller's HIP stream; no entry point blocks
 * the host except yee_energy_fp64 (which returns host scalars).
 *
 * Return codes: 0 ok, YEE_EINVAL bad argument, YEE_EHIP HIP error,
 * YEE_EUNSUPPORTED unsupported configuration; yee_last_error() gives a
 * thread-local message.
 */
#ifndef JAXPHYS_AMD_YEE_H
#define JAXPHYS_AMD_YEE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YEE_OK 0
#define YEE_EINVAL (-1)
#define YEE_EHIP (-2)
#define YEE_EUNSUPPORTED (-3)

/* Boundary treatment (fdtd3d.py:247-366).
 *  WALLS    : the reference's non-periodic sliced update with no damping
 *             (any boundary string other than "absorbing"/"periodic").
 *  SPONGE   : "absorbing": WALLS + E *= exp(-sigma_avg*dt/eps0) on every
 *             cell, sigma_avg from three 1-D cubic profiles (fdtd3d.py:139-167,
 *             221, 361-366).
 *  PERIODIC : "periodic": rolled curls on every cell, no damping
 *             (fdtd3d.py:247-301).
 *  CPML     : extension: WALLS + convolutional PML psi variables stored on
 *             the 2*npml-cell slabs of each axis only.                      */
#define YEE_BOUNDARY_WALLS 0
#define YEE_BOUNDARY_SPONGE 1
#define YEE_BOUNDARY_PERIODIC 2
#define YEE_BOUNDARY_CPML 3

/* Field component order used by every entry point. */
#define YEE_EX 0
#define YEE_EY 1
#define YEE_EZ 2
#define YEE_HX 3
#define YEE_HY 4
#define YEE_HZ 5

typedef struct YeeGridDesc {
  int32_t nx, ny, nz;   /* cells per axis, each >= 2                        */
  int32_t boundary;     /* YEE_BOUNDARY_*                                   */
  float ch;             /* dt / (mu0 dx)         fdtd3d.py:223               */
  const float* cb;      /* per cell: ce / eps_r  fdtd3d.py:224, 336-354      */
  const float* ca;      /* per cell E self-coefficient, NULL = 1 (lossless)  */
  /* SPONGE: sigma_x[nx], sigma_y[ny], sigma_z[nz] concatenated, plus the
   * float32-rounded dt and eps0 of fdtd3d.py:363.                         */
  const float* sponge_sigma;
  float sponge_dt;
  float sponge_eps0;
  /* CPML: slab thickness (same on all six faces, 2*npml <= n per axis) and
   * the coefficient tables: for axis a in x,y,z, for node kind E (integer
   * position) then H (half-integer position): b[n_a], c[n_a], 1/kappa[n_a];
   * i.e. 6*n_a floats per axis, axes concatenated x, y, z.                 */
  int32_t npml;
  const float* cpml;
  /* Optional material palette (no reference counterpart): when `mat` is not
   * NULL, mat[cell] (uint8) indexes `palette` = cb[256] then ca[256] (the
   * ca half is read only when `ca` is set) holding exactly the values of the
   * cb / ca arrays, which stay required.  The fused variants 3-4 and the
   * two-pass E kernel then read 1 byte per cell instead of 4 or 8.         */
  const uint8_t* mat;
  const float* palette;
} YeeGridDesc;

typedef struct YeePlan YeePlan;

/* Number of float32 values of the CPML psi buffer for this grid (0 unless
 * CPML): 12 arrays, each 2*npml cells thick along its derivative axis.  */
int64_t yee_psi_floats(const YeeGridDesc* desc);

/* Validate `desc` and build a plan (host object; its only device memory is
 * the source table uploaded by yee_set_sources).  Replaces the prologue of
 * fdtd3d.py:185-231 below the Python layer.                              */
int yee_plan_create(const YeeGridDesc* desc, YeePlan** out);
void yee_plan_destroy(YeePlan* plan);

/* Launch tuning (no reference counterpart): z planes marched per block by
 * the two-pass kernels (kz > 0) and by the fused kernel (fused_kz > 0;
 * 0 = automatic, < 0 = unchanged).  Results do not depend on them.       */
int yee_plan_tune(YeePlan* plan, int32_t kz, int32_t fused_kz);
/* Fused-kernel design: 0 = LDS exchange with a halo row, 1 = barrier-free
 * (H of column x-1 and row y-1 recomputed), 2 = barrier-free with
 * straight-line planes, carried E values and, under CPML, separate
 * launches for the slab-free interior and the slab shell, 3 (default) = 2
 * with the x neighbours exchanged between lanes.  Same results bit for
 * bit.                                                                   */
int yee_plan_set_variant(YeePlan* plan, int32_t variant);
/* Workgroup -> tile order (no reference counterpart): 0 = plain x-fastest
 * order, 1 = XCD-contiguous (each of the 8 XCDs runs a contiguous range of
 * tiles, so neighbouring tiles share an L2), 2 = XCD-banded (fused
 * interior launch: each z-chunk band split into 8 contiguous shares, one
 * per XCD; other launches use 1 for the two-pass and adjoint kernels, 0
 * for the fused shell), 3..10 = XCD row runs (each XCD takes runs of
 * order-2 consecutive rows of tiles while the 8 XCDs work on adjacent
 * runs), -1 (default) = automatic (1 for the two-pass and adjoint kernels,
 * 3 for the fused kernels).  Results do not depend on it.               */
int yee_plan_set_tile_order(YeePlan* plan, int32_t xcd_aware);
/* Replace the palette of the plan (both NULL: per-cell arrays only), e.g.
 * after the cb / ca arrays were changed in place.                        */
int yee_plan_set_palette(YeePlan* plan, const uint8_t* mat, const float* palette);
/* y rows (wavefronts) per workgroup tile of fused variants 2 and 3: 4
 * (default), 8 or 16 (no reference counterpart; results do not depend on
 * it).                                                                   */
int yee_plan_set_fused_rows(YeePlan* plan, int32_t rows);

#define YEE_FUSED_SOURCES 8

/* Point sources (fdtd3d.py:233-245, 368-369): source s adds table[n*nsrc+s]
 * to component comp[s] (YEE_EX/EY/EZ) at cell pos[3s..3s+2] = (i,j,k) after
 * the E update (and after sponge damping) of step n, in insertion order.
 * comp, pos and table are HOST arrays; the plan copies them.             */
int yee_set_sources(YeePlan* plan, int32_t nsrc, const int32_t* comp,
                    const int32_t* pos, const float* table, int32_t nsteps);

/* Run steps n0 .. n0+nsteps-1 (fdtd3d.py:280-393: H half-step with the old
 * E, then E half-step with the new H, damping, sources).  fields[6] in the
 * YEE_EX..YEE_HZ order; psi may be NULL unless CPML.                     */
int yee_run(YeePlan* plan, float* const* fields, float* psi, int32_t n0,
            int32_t nsteps, void* stream);

/* Fused single-pass step n (same arithmetic, bit-for-bit the same result as
 * yee_step_h + yee_step_e): reads fields_in / psi_in, writes fields_out /
 * psi_out (two distinct buffer sets).  One launch per step; 52 B of field
 * traffic per cell-update (56 B with Ca) instead of 76 / 80 B.           */
int yee_step_fused(YeePlan* plan, float* const* fields_in, float* const* fields_out,
                   float* psi_in, float* psi_out, int32_t n, void* stream);
/* Steps n0 .. n0+nsteps-1 alternating A->B, B->A, ...; the state ends in A
 * when nsteps is even, in B when odd.                                     */
int yee_run_fused(YeePlan* plan, float* const* fields_a, float* const* fields_b,
                  float* psi_a, float* psi_b, int32_t n0, int32_t nsteps, void* stream);

/* The two half-steps of step n separately (same semantics as yee_run). */
int yee_step_h(YeePlan* plan, float* const* fields, float* psi, void* stream);
int yee_step_e(YeePlan* plan, float* const* fields, float* psi, int32_t n,
               void* stream);

/* ---- adjoint (no reference counterpart) --------------------------------
 * The reference's gradient path is generic jax.grad through the scan
 * (src/neurosim/optimize.py:84; fdtd3d.py:393), which stores every step.
 * yee_adjoint_step applies the exact transpose of one forward step
 * (yee_step_h + yee_step_e, including sponge damping, CPML psi and lossy
 * Ca) to the adjoint state: lam[6] holds dJ/d(E^{n+1}, H^{n+1/2}) on entry
 * and dJ/d(E^n, H^{n-1/2}) on exit; lam_psi (same size as psi) holds the
 * psi adjoints, updated with one step of deferral for the H part (the
 * caller runs steps n = N-1 .. 0 in order and does nothing else with it):
 * each step first applies the psi_H-adjoint update left over from step n+1,
 * computed from the H part of lam on entry, so a sweep starts at step N
 * with the H part of lam zero (as it is for objectives over E) and lam_psi
 * zero.  With that start the sweep is the exact transpose of the forward
 * run: <lam_N, S^N x> = <lam_0, x> for states x with zero psi
 * (tests/test_adjoint_gpu.py).
 * Objective terms are added between steps with yee_probe_scatter_add.
 * Optional design box: cells where dJ/d(eps_r) is wanted, all with damping
 * 1 (outside sponge/CPML slabs) and every E component updated; the caller
 * passes d = E^{n+1} - Ca E^n of the forward step on the box (comp-major,
 * x fastest; what yee_run_record records) and, when acc_ca is given, E^n;
 * the kernel accumulates acc_cb += sum_c g_c d_c, acc_ca += sum_c g_c
 * E^n_c, with g = lam(E^{n+1}); then dJ/dCb = acc_cb / Cb and dJ/dCa =
 * acc_ca.                                                                 */
typedef struct YeeDesignBox {
  int32_t x0, y0, z0, nx, ny, nz;
  const float* d;
  const float* e_cur; /* E^n: needed only with acc_ca */
  float* acc_cb;
  float* acc_ca; /* may be NULL when lossless */
} YeeDesignBox;

int yee_adjoint_step(YeePlan* plan, float* const* lam, float* lam_psi,
                     const YeeDesignBox* box, void* stream);

/* Probe helpers for objectives (no reference counterpart):
 * gather  out[q] = fields[comp[q]][idx[q]]
 * scatter fields[comp[q]][idx[q]] += scale * w[q] * vals[q]  (w may be NULL)
 * comp (0..5, int32) and idx (cell index, int64) are device arrays; the
 * scatter indices must be distinct.                                      */
int yee_probe_gather(float* const* fields, const int32_t* comp, const int64_t* idx,
                     int32_t n, float* out, void* stream);
int yee_probe_scatter_add(float* const* fields, const int32_t* comp, const int64_t* idx,
                          const float* w, const float* vals, float scale, int32_t n,
                          void* stream);
/* Copy Ex, Ey, Ez on box = {x0, y0, z0, nx, ny, nz} (host array) into out
 * (comp-major, x fastest, 3*nx*ny*nz floats).                            */
int yee_box_gather(float* const* fields, int32_t nx, int32_t ny, int32_t nz,
                   const int32_t* box, float* out, void* stream);

/* yee_run plus recording after every step s = 0..nsteps-1 (step n = n0+s):
 * probe_out[s*nprobe + q] = field value of probe q after the step (as
 * yee_probe_gather), and the step's box record at box_out[s*R ...] with R =
 * 3*boxcells (R = 6*boxcells for lossy plans): d = E^{n+1} - Ca E^n on the
 * box (comp-major, x fastest), followed for lossy plans by E^n — what the
 * adjoint's eps gradient uses.  Either part may be disabled (nprobe == 0 /
 * box == NULL).  probe_box / probe_map
 * (optional, both or neither): the cell -> probe map of yee_adjoint_run,
 * for E probes at distinct (component, cell) pairs.  When the probes are
 * mapped (or absent) and every source is fused into the E kernel, that
 * kernel writes the records itself; otherwise one record launch follows
 * each step.  The recorded values are the same either way.              */
int yee_run_record(YeePlan* plan, float* const* fields, float* psi, int32_t n0,
                   int32_t nsteps, int32_t nprobe, const int32_t* probe_comp,
                   const int64_t* probe_idx, float* probe_out, const int32_t* box,
                   float* box_out, const int32_t* probe_box, const int32_t* probe_map,
                   void* stream);

/* Adjoint sweep over steps n_hi-1 down to n_lo (yee_adjoint_step each),
 * adding after step n (for n >= 1) the objective term
 * scale * w[q] * probe_hist[(n-1)*nprobe + q] to the adjoint of probe q
 * (the objective sum_n sum_q w_q F_q(after step n)^2 gives scale = 2).
 * Probes are E components at distinct (component, cell) pairs (merge
 * duplicates by adding their weights).
 * box history: the box records of steps n_lo .. n_hi-1 as yee_run_record
 * writes them (record m of step n_lo+m, R floats each).
 * probe_box / probe_map (optional, both or neither): a box (x0, y0, z0,
 * nx, ny, nz; host array) holding every probe cell and, on the device,
 * probe_map[((c*nz + k)*ny + j)*nx + i] = q for probe q of E component c
 * at (x0+i, y0+j, z0+k), -1 elsewhere.  With it the E-adjoint kernel adds
 * the objective terms itself (same values, no scatter launch per step).  */
int yee_adjoint_run(YeePlan* plan, float* const* lam, float* lam_psi, int32_t n_lo,
                    int32_t n_hi, int32_t nprobe, const int32_t* probe_comp,
                    const int64_t* probe_idx, const float* w, const float* probe_hist,
                    float scale, const int32_t* box, const float* box_hist,
                    float* acc_cb, float* acc_ca, const int32_t* probe_box,
                    const int32_t* probe_map, void* stream);

/* Float64 accumulators for the checks (no reference counterpart):
 * work[0] = sum_cells eps_r |E|^2 (eps_r from an explicit per-cell array,
 * NULL = 1), work[1] = sum_cells |H|^2.  `work` is a caller-owned device
 * buffer of YEE_ENERGY_WORK doubles; the summation order is fixed, so the
 * result is deterministic.  Asynchronous on the stream.                  */
#define YEE_ENERGY_WORK (2 + 2 * 1024)
int yee_energy_fp64(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                    const float* eps_r, double* work, void* stream);

/* yee_energy_fp64 for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Same kernels and the same
 * summation order as yee_energy_fp64, so the same bits; the 2*1024 float64
 * partial sums live in a static device array of the library (16 KiB per
 * GPU, held until the process ends) instead of a caller buffer, and `out`
 * (device, 2 doubles) receives sum eps_r |E|^2 and sum |H|^2.  Calls no
 * hipMalloc/hipFree.  Grids above YEE_NOALLOC_MAX_CELLS cells are rejected
 * with YEE_EINVAL.  All calls on one GPU share the array: calls that
 * overlap in time on one GPU (different streams or host threads) race;
 * the caller's stream must belong to the current device.                 */
#define YEE_NOALLOC_MAX_CELLS (1LL << 36)
int yee_energy_fp64_noalloc(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                            const float* eps_r, double* out, void* stream);

/* Float64 Poynting flux through the plane normal to `axis` (0 x, 1 y, 2 z)
 * halfway between the E planes `plane` and `plane+1`, where the tangential
 * H components live: work[0] = sum over the plane of (E x H) . n, each
 * tangential E the mean of its two planes (no reference counterpart; for
 * energy-balance checks and field diagnostics, dx^2 times it is the power
 * through the plane).  `work` holds YEE_FLUX_WORK doubles; deterministic.  */
#define YEE_FLUX_WORK (1 + 1024)
int yee_poynting_flux_fp64(int32_t nx, int32_t ny, int32_t nz, float* const* fields,
                           int32_t axis, int32_t plane, double* work, void* stream);

/* Plain device copy of n floats, used to calibrate HBM byte counters
 * (no reference counterpart).                                            */
int yee_copy_probe(const float* src, float* dst, int64_t n, void* stream);

const char* yee_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* JAXPHYS_AMD_YEE_H */
