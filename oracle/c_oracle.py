"""ctypes wrapper of the C/OpenMP oracle (liboracle_fdtd.so).

TEST INFRASTRUCTURE ONLY: the checker for tests/ at sizes where NumPy is
slow, the float64 error oracle, and bench.py's cpu_baseline leg.  It computes
its own coefficients from ``oracle.ext_fdtd`` / ``oracle.ref_fdtd`` (never
from the product) and works in the reference layout at its interface
(arrays ``(nx, ny, nz)``), x fastest inside.
"""

from __future__ import annotations

import ctypes
import os
import time

import numpy as np

from oracle import ext_fdtd as X
from oracle import ref_fdtd as R

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_fdtd.so")
BOUNDARY = {"walls": 0, "absorbing": 1, "periodic": 2, "cpml": 3}


class OracleDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("boundary", ctypes.c_int32), ("npml", ctypes.c_int32),
                ("ch", ctypes.c_double),
                ("cb", ctypes.c_void_p), ("ca", ctypes.c_void_p), ("damp", ctypes.c_void_p),
                ("cpml", ctypes.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from oracle import build_oracle
            build_oracle.build()
        h = ctypes.CDLL(LIB_PATH)
        for nm in ("oracle_run_f32", "oracle_run_f64"):
            fn = getattr(h, nm)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.POINTER(OracleDesc), ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                           ctypes.c_int, ctypes.c_int, ctypes.c_int]
        h.oracle_psi_count.restype = ctypes.c_longlong
        h.oracle_psi_count.argtypes = [ctypes.POINTER(OracleDesc)]
        h.oracle_max_threads.restype = ctypes.c_int
        _lib = h
    return _lib


def host_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def _internal(a, dtype):
    """(nx, ny, nz) -> contiguous x-fastest storage (nz, ny, nx)."""
    return np.ascontiguousarray(np.asarray(a, dtype=dtype).transpose(2, 1, 0))


class COracle:
    """State of one oracle run: fields (x fastest), psi, coefficients."""

    def __init__(self, size, dx, dt, boundary, eps_r, sigma=None, npml=0, cpml_kw=None,
                 dtype=np.float64, sponge_layers=None, cpml_coeffs=None):
        """``cpml_coeffs`` (tests only) replaces ``ext_fdtd.cpml_coeffs`` as
        the function that builds each axis's CPML tables, e.g. to plant a
        mistake and show that a test catches it."""
        self.dtype = np.dtype(dtype).type
        self.size = tuple(size)
        nx, ny, nz = size
        ca, cb = X.e_coeffs(eps_r, sigma, dt, dx, self.dtype)
        self.cb = _internal(cb, self.dtype)
        self.ca = _internal(ca, self.dtype) if not np.all(ca == 1) else None
        self.damp = None
        if boundary == "absorbing":
            lay = npml if sponge_layers is None else sponge_layers
            self.damp = _internal(R.damping_factor(R.sigma_avg(size, lay, dx, self.dtype), dt,
                                                   self.dtype), self.dtype)
        self.cp = None
        if boundary == "cpml":
            tabs = []
            fn = X.cpml_coeffs if cpml_coeffs is None else cpml_coeffs
            for n in size:
                t = fn(n, npml, dx, dt, **(cpml_kw or {}))
                tabs.append(np.concatenate([np.asarray(v) for v in (*t["e"], *t["h"])]))
            self.cp = np.ascontiguousarray(np.concatenate(tabs).astype(self.dtype))
        d = OracleDesc()
        d.nx, d.ny, d.nz = nx, ny, nz
        d.boundary = BOUNDARY.get(boundary, 0)
        d.npml = npml if boundary == "cpml" else 0
        d.ch = float(self.dtype(dt / (float(R.MU0) * dx)))
        d.cb = self.cb.ctypes.data
        d.ca = self.ca.ctypes.data if self.ca is not None else None
        d.damp = self.damp.ctypes.data if self.damp is not None else None
        d.cpml = self.cp.ctypes.data if self.cp is not None else None
        self.desc = d
        shape = (nz, ny, nx)
        self.f = [np.zeros(shape, dtype=self.dtype) for _ in range(6)]
        npsi = int(lib().oracle_psi_count(ctypes.byref(d)))
        self.psi = np.zeros(max(npsi, 1), dtype=self.dtype)
        self.set_sources([])

    def set_sources(self, sources):
        """``sources``: ``[(comp, (i, j, k), table)]`` with per-step values."""
        nx, ny, _ = self.size
        self.nsrc = len(sources)
        self.comp = np.asarray([c for c, _, _ in sources], dtype=np.int32)
        self.sidx = np.asarray([(k * ny + j) * nx + i for _, (i, j, k), _ in sources],
                               dtype=np.int64)
        if sources:
            self.table = np.ascontiguousarray(
                np.stack([np.asarray(t, dtype=np.float64) for _, _, t in sources], axis=1)
                .astype(self.dtype))
        else:
            self.table = np.zeros((1, 1), dtype=self.dtype)

    def run(self, n0, nsteps, threads=0):
        if self.nsrc and n0 + nsteps > self.table.shape[0]:
            raise ValueError("source table too short")
        fp = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in self.f])
        fn = lib().oracle_run_f32 if self.dtype == np.float32 else lib().oracle_run_f64
        rc = fn(ctypes.byref(self.desc), ctypes.cast(fp, ctypes.c_void_p), self.psi.ctypes.data,
                self.nsrc, self.comp.ctypes.data, self.sidx.ctypes.data, self.table.ctypes.data,
                int(n0), int(nsteps), int(threads))
        assert rc == 0

    def fields_ref(self):
        """Fields in the reference's (nx, ny, nz) order."""
        return [a.transpose(2, 1, 0) for a in self.f]


def time_arrays(size, cb, ca, cpml_tabs, npml, ch, sources_table, comps, sidx, budget_s,
                threads=0):
    """Time float32 steps of the C oracle on given coefficient arrays (x
    fastest), for about ``budget_s`` seconds; returns (steps, seconds)."""
    nx, ny, nz = size
    d = OracleDesc()
    d.nx, d.ny, d.nz = nx, ny, nz
    d.boundary = 3 if cpml_tabs is not None else 0
    d.npml = npml if cpml_tabs is not None else 0
    d.ch = float(ch)
    d.cb = cb.ctypes.data
    d.ca = ca.ctypes.data if ca is not None else None
    d.damp = None
    d.cpml = cpml_tabs.ctypes.data if cpml_tabs is not None else None
    shape = (nz, ny, nx)
    f = [np.zeros(shape, np.float32) for _ in range(6)]
    npsi = int(lib().oracle_psi_count(ctypes.byref(d)))
    psi = np.zeros(max(npsi, 1), np.float32)
    fp = (ctypes.c_void_p * 6)(*[a.ctypes.data for a in f])
    comps = np.ascontiguousarray(comps, np.int32)
    sidx = np.ascontiguousarray(sidx, np.int64)
    tab = np.ascontiguousarray(sources_table, np.float32)

    def go(n0, n):
        t = time.perf_counter()
        lib().oracle_run_f32(ctypes.byref(d), ctypes.cast(fp, ctypes.c_void_p), psi.ctypes.data,
                             len(comps), comps.ctypes.data, sidx.ctypes.data, tab.ctypes.data,
                             n0, n, threads)
        return time.perf_counter() - t

    go(0, 1)  # first touch of the arrays
    t2 = go(1, 2) / 2.0
    n = max(1, min(tab.shape[0] - 3, int(budget_s / max(t2, 1e-6))))
    return n, go(3, n)
