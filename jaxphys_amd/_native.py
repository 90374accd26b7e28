"""ctypes binding of libyee_hip.so (C ABI in include/yee.h).

The library is loaded after ``import torch`` so that it binds to the HIP
runtime torch already loaded (both name it libamdhip64.so.7): device
pointers of torch tensors and torch's streams are then valid inside it.
There is no fallback: if the library is missing or no HIP device is present
every call raises ``NativeLibraryError``.
"""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (must precede the library load, see above)

from jaxphys_amd.exceptions import ConfigurationError, NativeLibraryError, SimulationError

HERE = os.path.dirname(os.path.abspath(__file__))

YEE_OK, YEE_EINVAL, YEE_EHIP, YEE_EUNSUPPORTED = 0, -1, -2, -3
BOUNDARY_WALLS, BOUNDARY_SPONGE, BOUNDARY_PERIODIC, BOUNDARY_CPML = 0, 1, 2, 3
FUSED_SOURCES = 8
ENERGY_WORK = 2 + 2 * 1024

_c_float_p = ctypes.c_void_p


class YeeGridDesc(ctypes.Structure):
    """Mirror of ``YeeGridDesc`` in include/yee.h."""

    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("boundary", ctypes.c_int32),
        ("ch", ctypes.c_float),
        ("cb", ctypes.c_void_p), ("ca", ctypes.c_void_p),
        ("sponge_sigma", ctypes.c_void_p),
        ("sponge_dt", ctypes.c_float), ("sponge_eps0", ctypes.c_float),
        ("npml", ctypes.c_int32),
        ("cpml", ctypes.c_void_p),
        ("mat", ctypes.c_void_p), ("palette", ctypes.c_void_p),
    ]


_LIB = None

EXPORTS = {
    "yee_psi_floats": (ctypes.c_int64, [ctypes.POINTER(YeeGridDesc)]),
    "yee_plan_create": (ctypes.c_int, [ctypes.POINTER(YeeGridDesc), ctypes.POINTER(ctypes.c_void_p)]),
    "yee_plan_destroy": (None, [ctypes.c_void_p]),
    "yee_plan_tune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32]),
    "yee_plan_set_variant": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "yee_plan_set_tile_order": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "yee_plan_set_fused_rows": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "yee_plan_set_palette": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "yee_set_sources": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]),
    "yee_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p]),
    "yee_step_h": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "yee_step_e": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_int32, ctypes.c_void_p]),
    "yee_energy_fp64": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p]),
    "yee_energy_fp64_noalloc": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p]),
    "yee_copy_probe": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                      ctypes.c_void_p]),
    "yee_poynting_flux_fp64": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p]),
    "yee_last_error": (ctypes.c_char_p, []),
    "yee_step_fused": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                      ctypes.c_void_p]),
    "yee_run_fused": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                     ctypes.c_int32, ctypes.c_void_p]),
    "yee_adjoint_step": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_void_p, ctypes.c_void_p]),
    "yee_probe_gather": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "yee_probe_scatter_add": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_float,
                                             ctypes.c_int32, ctypes.c_void_p]),
    "yee_box_gather": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p]),
    "yee_run_record": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                      ctypes.c_void_p, ctypes.c_void_p]),
    "yee_adjoint_run": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_float, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}


class YeeDesignBox(ctypes.Structure):
    """Mirror of ``YeeDesignBox`` in include/yee.h."""

    _fields_ = [("x0", ctypes.c_int32), ("y0", ctypes.c_int32), ("z0", ctypes.c_int32),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("d", ctypes.c_void_p), ("e_cur", ctypes.c_void_p),
                ("acc_cb", ctypes.c_void_p), ("acc_ca", ctypes.c_void_p)]


def library_path(name: str = "libyee_hip.so") -> str:
    return os.path.join(HERE, name)


def lib():
    """Load (once) and return the HIP library; raise if it is not built."""
    global _LIB
    if _LIB is None:
        path = library_path()
        if not os.path.exists(path):
            raise NativeLibraryError(
                f"{path} is not built; run `python -m jaxphys_amd.build` "
                "(hipcc, gfx950) first")
        handle = ctypes.CDLL(path)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = handle
    return _LIB


def require_device() -> torch.device:
    if not torch.cuda.is_available():
        raise NativeLibraryError(
            "jaxphys_amd runs its FDTD step on an AMD GPU (HIP); no device is visible")
    return torch.device("cuda", torch.cuda.current_device())


def check(rc: int, what: str) -> None:
    if rc == YEE_OK:
        return
    msg = lib().yee_last_error().decode(errors="replace")
    if rc in (YEE_EINVAL, YEE_EUNSUPPORTED):
        raise ConfigurationError(f"{what}: {msg}")
    raise SimulationError(f"{what}: {msg}")


def ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def field_array(fields) -> ctypes.Array:
    return (ctypes.c_void_p * 6)(*[ptr(f) for f in fields])


def stream_handle(stream=None) -> int:
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)


class Plan:
    """Owns a ``YeePlan*``; the tensors named in the descriptor must outlive it."""

    def __init__(self, desc: YeeGridDesc, keep=()):
        self.desc = desc
        self._keep = list(keep)  # tensors whose pointers the descriptor holds
        h = ctypes.c_void_p()
        check(lib().yee_plan_create(ctypes.byref(desc), ctypes.byref(h)), "yee_plan_create")
        self.handle = h

    def tune(self, kz: int = 0, fused_kz: int = -1) -> None:
        check(lib().yee_plan_tune(self.handle, int(kz), int(fused_kz)), "yee_plan_tune")

    def set_variant(self, variant: int) -> None:
        check(lib().yee_plan_set_variant(self.handle, int(variant)), "yee_plan_set_variant")

    def set_tile_order(self, order) -> None:
        """None: automatic (default); False / 0: plain; True / 1:
        XCD-contiguous; 2: XCD-banded; 3 .. 10: XCD runs of order - 2
        rows of tiles."""
        v = -1 if order is None else int(order)
        check(lib().yee_plan_set_tile_order(self.handle, v), "yee_plan_set_tile_order")

    def set_palette(self, mat, palette) -> None:
        """uint8 material ids per cell and the 512-float (cb | ca) table, or
        None, None for the per-cell arrays only."""
        check(lib().yee_plan_set_palette(self.handle, None if mat is None else ptr(mat),
                                         None if palette is None else ptr(palette)),
              "yee_plan_set_palette")
        self._palette_refs = (mat, palette)

    def set_fused_rows(self, rows: int) -> None:
        check(lib().yee_plan_set_fused_rows(self.handle, int(rows)), "yee_plan_set_fused_rows")

    def psi_floats(self) -> int:
        return int(lib().yee_psi_floats(ctypes.byref(self.desc)))

    def set_sources(self, comp, pos, table, nsteps: int) -> None:
        import numpy as np
        comp = np.ascontiguousarray(comp, dtype=np.int32)
        pos = np.ascontiguousarray(pos, dtype=np.int32).reshape(-1)
        table = np.ascontiguousarray(table, dtype=np.float32).reshape(-1)
        nsrc = int(comp.shape[0])
        check(lib().yee_set_sources(self.handle, nsrc, comp.ctypes.data, pos.ctypes.data,
                                    table.ctypes.data if table.size else None, int(nsteps)),
              "yee_set_sources")

    def run(self, fields, psi, n0: int, nsteps: int, stream=None) -> None:
        check(lib().yee_run(self.handle, field_array(fields), ptr(psi) or None, int(n0),
                            int(nsteps), stream_handle(stream)), "yee_run")

    def step_h(self, fields, psi, stream=None) -> None:
        check(lib().yee_step_h(self.handle, field_array(fields), ptr(psi) or None,
                               stream_handle(stream)), "yee_step_h")

    def step_e(self, fields, psi, n: int, stream=None) -> None:
        check(lib().yee_step_e(self.handle, field_array(fields), ptr(psi) or None, int(n),
                               stream_handle(stream)), "yee_step_e")

    def close(self) -> None:
        if getattr(self, "handle", None) is not None and self.handle.value:
            lib().yee_plan_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_fused(plan: "Plan", fa, fb, psi_a, psi_b, n0, nsteps, stream=None) -> None:
    check(lib().yee_run_fused(plan.handle, field_array(fa), field_array(fb), ptr(psi_a) or None,
                              ptr(psi_b) or None, int(n0), int(nsteps), stream_handle(stream)),
          "yee_run_fused")


def step_fused(plan: "Plan", fin, fout, psi_in, psi_out, n, stream=None) -> None:
    check(lib().yee_step_fused(plan.handle, field_array(fin), field_array(fout),
                               ptr(psi_in) or None, ptr(psi_out) or None, int(n),
                               stream_handle(stream)), "yee_step_fused")


def box_array(box):
    return None if box is None else (ctypes.c_int32 * 6)(*[int(v) for v in box])


def run_record(plan: "Plan", fields, psi, n0, nsteps, probe_comp=None, probe_idx=None,
               probe_out=None, box=None, box_out=None, probe_box=None, probe_map=None,
               stream=None) -> None:
    """``yee_run_record``: steps with per-step probe / design-box recording."""
    nprobe = 0 if probe_comp is None else int(probe_comp.numel())
    check(lib().yee_run_record(plan.handle, field_array(fields), ptr(psi) or None, int(n0),
                               int(nsteps), nprobe, ptr(probe_comp) or None,
                               ptr(probe_idx) or None, ptr(probe_out) or None,
                               box_array(box), ptr(box_out) or None, box_array(probe_box),
                               ptr(probe_map) or None, stream_handle(stream)),
          "yee_run_record")


def adjoint_run(plan: "Plan", lam, lam_psi, n_lo, n_hi, probe_comp, probe_idx, w, probe_hist,
                scale, box=None, box_hist=None, acc_cb=None, acc_ca=None, probe_box=None,
                probe_map=None, stream=None) -> None:
    nprobe = 0 if probe_comp is None else int(probe_comp.numel())
    check(lib().yee_adjoint_run(plan.handle, field_array(lam), ptr(lam_psi) or None, int(n_lo),
                                int(n_hi), nprobe, ptr(probe_comp) or None,
                                ptr(probe_idx) or None, ptr(w) or None, ptr(probe_hist) or None,
                                float(scale), box_array(box), ptr(box_hist) or None,
                                ptr(acc_cb) or None, ptr(acc_ca) or None, box_array(probe_box),
                                ptr(probe_map) or None, stream_handle(stream)),
          "yee_adjoint_run")


def probe_gather(fields, comp, idx, out, stream=None) -> None:
    check(lib().yee_probe_gather(field_array(fields), ptr(comp), ptr(idx), int(comp.numel()),
                                 ptr(out), stream_handle(stream)), "yee_probe_gather")


def probe_scatter_add(fields, comp, idx, w, vals, scale, stream=None) -> None:
    check(lib().yee_probe_scatter_add(field_array(fields), ptr(comp), ptr(idx), ptr(w) or None,
                                      ptr(vals), float(scale), int(comp.numel()),
                                      stream_handle(stream)), "yee_probe_scatter_add")


def box_gather(fields, shape_xyz, box, out, stream=None) -> None:
    nx, ny, nz = shape_xyz
    check(lib().yee_box_gather(field_array(fields), nx, ny, nz, box_array(box), ptr(out),
                               stream_handle(stream)), "yee_box_gather")


def adjoint_step(plan: "Plan", lam, lam_psi, box=None, stream=None) -> None:
    check(lib().yee_adjoint_step(plan.handle, field_array(lam), ptr(lam_psi) or None,
                                 ctypes.byref(box) if box is not None else None,
                                 stream_handle(stream)), "yee_adjoint_step")


def energy_fp64(shape_xyz, fields, eps_r=None, work=None, stream=None):
    """float64 (sum eps_r |E|^2, sum |H|^2) over the grid; returns a device
    tensor of 2 doubles (asynchronous)."""
    nx, ny, nz = shape_xyz
    dev = fields[0].device
    if work is None:
        work = torch.empty(ENERGY_WORK, dtype=torch.float64, device=dev)
    check(lib().yee_energy_fp64(nx, ny, nz, field_array(fields), ptr(eps_r) or None,
                                ptr(work), stream_handle(stream)), "yee_energy_fp64")
    return work[:2]


def energy_fp64_noalloc(shape_xyz, fields, eps_r, out, stream=None) -> None:
    """``energy_fp64`` through the entry point that allocates nothing: the
    partial sums stay in the library's static device array; ``out`` (device
    float64, 2 values) receives (sum eps_r |E|^2, sum |H|^2).  Asynchronous."""
    nx, ny, nz = shape_xyz
    check(lib().yee_energy_fp64_noalloc(nx, ny, nz, field_array(fields), ptr(eps_r) or None,
                                        ptr(out), stream_handle(stream)),
          "yee_energy_fp64_noalloc")


FLUX_WORK = 1 + 1024


def poynting_flux_fp64(shape_xyz, fields, axis: int, plane: int, work=None, stream=None):
    """float64 sum over a Yee H plane of (E x H) . n (see include/yee.h);
    returns a device tensor of one double (asynchronous)."""
    nx, ny, nz = shape_xyz
    if work is None:
        work = torch.empty(FLUX_WORK, dtype=torch.float64, device=fields[0].device)
    check(lib().yee_poynting_flux_fp64(nx, ny, nz, field_array(fields), int(axis), int(plane),
                                       ptr(work), stream_handle(stream)),
          "yee_poynting_flux_fp64")
    return work[:1]


def copy_probe(src: torch.Tensor, dst: torch.Tensor, stream=None) -> None:
    check(lib().yee_copy_probe(ptr(src), ptr(dst), int(src.numel()), stream_handle(stream)),
          "yee_copy_probe")
