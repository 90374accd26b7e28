/* C restatement of the Yee step, instantiated for float and double by
 * fdtd_oracle.c (T = element type, SFX = name suffix).  TEST
 * INFRASTRUCTURE: the CPU checker and CPU timing baseline, never the
 * product.
 *
 * Follows src/neurosim/em/fdtd3d.py:280-379 of the reference:
 *   H update ranges and formulas   fdtd3d.py:303-330 (periodic :290-295)
 *   E update ranges and formulas   fdtd3d.py:332-359 (periodic :297-301)
 *   sponge damping, E only         fdtd3d.py:361-366 (per-cell factor given)
 *   sources after damping, in order fdtd3d.py:236-244, 368-369
 * plus the CPML / lossy extensions restated in oracle/ext_fdtd.py.
 * Storage is x fastest: idx = (k*ny + j)*nx + i, i.e. reference F[i,j,k].
 * psi storage: slab-only, 12 arrays (see psi_slot below).                 */

#define CAT_(a, b) a##b
#define CAT(a, b) CAT_(a, b)

static inline T CAT(cterm, SFX)(T* ps, const T* tab, long n, long a, int kind, T d) {
  T b = tab[(3 * kind + 0) * n + a], c = tab[(3 * kind + 1) * n + a],
    kinv = tab[(3 * kind + 2) * n + a];
  T v = b * (*ps) + c * d;
  *ps = v;
  return d * kinv + v;
}

int CAT(oracle_run_, SFX)(const OracleDesc* o, T* const* f, T* psi, int nsrc, const int* comp,
                          const long long* sidx, const T* table, int n0, int nsteps,
                          int nthreads) {
  const long nx = o->nx, ny = o->ny, nz = o->nz, P = nx * ny;
  const int per = o->boundary == OR_PERIODIC, cp = o->boundary == OR_CPML;
  const long p = cp ? o->npml : 0;
  const T ch = (T)o->ch;
  const T* cb = (const T*)o->cb;
  const T* ca = (const T*)o->ca;
  const T* damp = (const T*)o->damp;
  const T* tx = (const T*)o->cpml;
  const T* ty = tx ? tx + 6 * nx : 0;
  const T* tz = tx ? tx + 6 * (nx + ny) : 0;
  T *ex = f[0], *ey = f[1], *ez = f[2], *hx = f[3], *hy = f[4], *hz = f[5];
  /* psi slabs: y-arrays (nz,2p,nx), z-arrays (2p,ny,nx), x-arrays (nz,ny,2p) */
  const long ysz = nz * 2 * p * nx, zsz = 2 * p * ny * nx, xsz = nz * ny * 2 * p;
  T *pExy = psi, *pEzy = psi + ysz, *pHxy = psi + 2 * ysz, *pHzy = psi + 3 * ysz;
  T *pExz = psi + 4 * ysz, *pEyz = pExz + zsz, *pHxz = pExz + 2 * zsz, *pHyz = pExz + 3 * zsz;
  T *pEyx = pExz + 4 * zsz, *pEzx = pEyx + xsz, *pHyx = pEyx + 2 * xsz, *pHzx = pEyx + 3 * xsz;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  for (int n = n0; n < n0 + nsteps; ++n) {
    /* ---- H half-step ---- */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nz; ++k) {
      const int zin = per || k < nz - 1;
      const long kp = (k + 1 == nz) ? 0 : k + 1;
      const int zs = cp && (k < p || k >= nz - p);
      const long sz = k < p ? k : k - (nz - 2 * p);
      for (long j = 0; j < ny; ++j) {
        const int yin = per || j < ny - 1;
        const long jp = (j + 1 == ny) ? 0 : j + 1;
        const int ys = cp && (j < p || j >= ny - p);
        const long sy = j < p ? j : j - (ny - 2 * p);
        for (long i = 0; i < nx; ++i) {
          const int xin = per || i < nx - 1;
          const long ip = (i + 1 == nx) ? 0 : i + 1;
          const int xs = cp && (i < p || i >= nx - p);
          const long sx = i < p ? i : i - (nx - 2 * p);
          const long c = k * P + j * nx + i;
          const long cI = k * P + j * nx + ip, cJ = k * P + jp * nx + i, cK = kp * P + j * nx + i;
          if (yin && zin) {
            T a = ez[cJ] - ez[c], b = ey[cK] - ey[c];
            if (ys) a = CAT(cterm, SFX)(pHxy + (k * 2 * p + sy) * nx + i, ty, ny, j, 1, a);
            if (zs) b = CAT(cterm, SFX)(pHxz + (sz * ny + j) * nx + i, tz, nz, k, 1, b);
            hx[c] = hx[c] - ch * (a - b);
          }
          if (xin && zin) {
            T a = ex[cK] - ex[c], b = ez[cI] - ez[c];
            if (zs) a = CAT(cterm, SFX)(pHyz + (sz * ny + j) * nx + i, tz, nz, k, 1, a);
            if (xs) b = CAT(cterm, SFX)(pHyx + (k * ny + j) * 2 * p + sx, tx, nx, i, 1, b);
            hy[c] = hy[c] - ch * (a - b);
          }
          if (xin && yin) {
            T a = ey[cI] - ey[c], b = ex[cJ] - ex[c];
            if (xs) a = CAT(cterm, SFX)(pHzx + (k * ny + j) * 2 * p + sx, tx, nx, i, 1, a);
            if (ys) b = CAT(cterm, SFX)(pHzy + (k * 2 * p + sy) * nx + i, ty, ny, j, 1, b);
            hz[c] = hz[c] - ch * (a - b);
          }
        }
      }
    }
    /* ---- E half-step ---- */
#pragma omp parallel for schedule(static)
    for (long k = 0; k < nz; ++k) {
      const int zin = per || k > 0;
      const long km = (k == 0) ? nz - 1 : k - 1;
      const int zs = cp && (k < p || k >= nz - p);
      const long sz = k < p ? k : k - (nz - 2 * p);
      for (long j = 0; j < ny; ++j) {
        const int yin = per || j > 0;
        const long jm = (j == 0) ? ny - 1 : j - 1;
        const int ys = cp && (j < p || j >= ny - p);
        const long sy = j < p ? j : j - (ny - 2 * p);
        for (long i = 0; i < nx; ++i) {
          const int xin = per || i > 0;
          const long im = (i == 0) ? nx - 1 : i - 1;
          const int xs = cp && (i < p || i >= nx - p);
          const long sx = i < p ? i : i - (nx - 2 * p);
          const long c = k * P + j * nx + i;
          const long cI = k * P + j * nx + im, cJ = k * P + jm * nx + i, cK = km * P + j * nx + i;
          const T b_ = cb[c];
          if (yin && zin) {
            T a = hz[c] - hz[cJ], b = hy[c] - hy[cK];
            if (ys) a = CAT(cterm, SFX)(pExy + (k * 2 * p + sy) * nx + i, ty, ny, j, 0, a);
            if (zs) b = CAT(cterm, SFX)(pExz + (sz * ny + j) * nx + i, tz, nz, k, 0, b);
            ex[c] = ca ? ca[c] * ex[c] + b_ * (a - b) : ex[c] + b_ * (a - b);
          }
          if (xin && zin) {
            T a = hx[c] - hx[cK], b = hz[c] - hz[cI];
            if (zs) a = CAT(cterm, SFX)(pEyz + (sz * ny + j) * nx + i, tz, nz, k, 0, a);
            if (xs) b = CAT(cterm, SFX)(pEyx + (k * ny + j) * 2 * p + sx, tx, nx, i, 0, b);
            ey[c] = ca ? ca[c] * ey[c] + b_ * (a - b) : ey[c] + b_ * (a - b);
          }
          if (xin && yin) {
            T a = hy[c] - hy[cI], b = hx[c] - hx[cJ];
            if (xs) a = CAT(cterm, SFX)(pEzx + (k * ny + j) * 2 * p + sx, tx, nx, i, 0, a);
            if (ys) b = CAT(cterm, SFX)(pEzy + (k * 2 * p + sy) * nx + i, ty, ny, j, 0, b);
            ez[c] = ca ? ca[c] * ez[c] + b_ * (a - b) : ez[c] + b_ * (a - b);
          }
          if (damp) {
            const T d = damp[c];
            ex[c] *= d;
            ey[c] *= d;
            ez[c] *= d;
          }
        }
      }
    }
    /* ---- sources ---- */
    for (int q = 0; q < nsrc; ++q) {
      T* g = comp[q] == 0 ? ex : (comp[q] == 1 ? ey : ez);
      g[sidx[q]] += table[(long)n * nsrc + q];
    }
  }
  return 0;
}

#undef CAT
#undef CAT_
