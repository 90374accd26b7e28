/*
 * jaxphys_amd -- C ABI of the MI355X (gfx950) 2D TM FDTD library
 * (libtm2d_hip.so).
 *
 * The reference's 2D solver (sushaan-k/jaxphys, package neurosim) is the
 * Python closure `step` inside `EMGrid.simulate`
 * (src/neurosim/em/fdtd.py:298-353) driven by `jax.lax.scan`
 * (fdtd.py:355-359); it has no FFI.  These entry points replace that scan:
 * the Python `EMGrid` (same API as fdtd.py:74-386) builds the tables once
 * and calls them through ctypes.
 *
 * Storage: Ez, Hx, Hy and every per-cell table are x fastest,
 *     idx(i, j) = j * nx + i,
 * where (i, j) are the reference's indices F[i, j] (fdtd.py:225-227).  All
 * data pointers are device pointers owned by the caller (torch tensors),
 * borrowed for the duration of a call; `fields` arguments are host arrays of
 * three device pointers in the order Ez, Hx, Hy.  Launches go on the
 * caller's HIP stream; no entry point blocks the host.
 *
 * Return codes: 0 ok, TM2D_EINVAL bad argument, TM2D_EHIP HIP error;
 * tm2d_last_error() gives a thread-local message.
 */
#ifndef JAXPHYS_AMD_TM2D_H
#define JAXPHYS_AMD_TM2D_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TM2D_OK 0
#define TM2D_EINVAL (-1)
#define TM2D_EHIP (-2)

/* fdtd.py:96-100, 337-347 (the boundary literals of config.py:108). */
#define TM2D_ABSORBING 0  /* sliced update, Ez *= exp(-sigma_avg dt/eps0)   */
#define TM2D_PERIODIC 1   /* rolled curls, then the periodic edge copies    */
#define TM2D_REFLECTING 2 /* sliced update, then the mirror edge copies     */

typedef struct Tm2dDesc {
  int32_t nx, ny;           /* cells per axis, each >= 3                     */
  int32_t boundary;         /* TM2D_*                                        */
  int32_t npml;             /* ABSORBING: cells with i or j within npml of an
                               edge are the only ones whose damping factor
                               differs from 1 (fdtd.py:176-189)              */
  float ch;                 /* dt / (mu0 dx)   fdtd.py:240                    */
  float ce;                 /* dt / (eps0 dx)  fdtd.py:241                    */
  const float* damp;        /* ABSORBING: per-cell exp(-sigma_avg dt/eps0)
                               (fdtd.py:229, 339), read in the band only     */
  const uint8_t* conductor; /* NULL, or per cell: nonzero = Ez forced to 0
                               after the sources (fdtd.py:143-157, 351)      */
  int32_t n_src;            /* plane-wave sources (fdtd.py:249-254)          */
  int32_t table_steps;      /* rows of src_table                             */
  const int32_t* src_rows;  /* device, n_src: the source rows y              */
  const float* src_table;   /* device, table_steps x n_src: the value added
                               to row y at step n (amplitude*sin(w t_n)),
                               sources in insertion order                    */
} Tm2dDesc;

/* One step n (fdtd.py:298-353): reads Ez^n, Hx, Hy from `in`, writes the
 * new fields to `out` (distinct buffers).  Replaces one iteration of the
 * scan at fdtd.py:357. */
int tm2d_step(const Tm2dDesc* d, float* const* in, float* const* out, int32_t n,
              void* stream);

/* Steps n0 .. n0+nsteps-1, ping-ponging between the buffer sets `a` (holds
 * the state at entry) and `b`: the result is in `a` when nsteps is even and
 * in `b` when it is odd.  Replaces the scan of fdtd.py:355-359. */
int tm2d_run(const Tm2dDesc* d, float* const* a, float* const* b, int32_t n0, int32_t nsteps,
             void* stream);

const char* tm2d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
