/* CPU oracle of the 3D Yee step in C/OpenMP (float32 and float64).
 * TEST INFRASTRUCTURE ONLY: used by tests/ as a checker at sizes the NumPy
 * restatement is too slow for, and by bench.py's cpu_baseline leg as the
 * CPU timing baseline.  The product never links or loads it.
 * Semantics: see fdtd_oracle_impl.h (reference src/neurosim/em/fdtd3d.py). */
#include <omp.h>
#include <stdint.h>

#define OR_WALLS 0
#define OR_SPONGE 1
#define OR_PERIODIC 2
#define OR_CPML 3

typedef struct OracleDesc {
  int32_t nx, ny, nz, boundary, npml;
  double ch;
  const void* cb;   /* per cell */
  const void* ca;   /* per cell or NULL */
  const void* damp; /* per cell or NULL (sponge) */
  const void* cpml; /* per axis 6*n: bE cE kE bH cH kH */
} OracleDesc;

long long oracle_psi_count(const OracleDesc* o) {
  if (o->boundary != OR_CPML) return 0;
  const long long p2 = 2LL * o->npml;
  return 4 * ((long long)o->nz * p2 * o->nx + p2 * o->ny * o->nx + (long long)o->nz * o->ny * p2);
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

#define T float
#define SFX f32
#include "fdtd_oracle_impl.h"
#undef T
#undef SFX

#define T double
#define SFX f64
#include "fdtd_oracle_impl.h"
#undef T
#undef SFX
