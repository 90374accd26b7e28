"""ctypes binding of libconv3d_hip.so (C ABI in include/conv3d.h)."""

from __future__ import annotations

import ctypes
import os

import torch  # noqa: F401  (load order: torch's HIP runtime first)

from jaxphys_amd._native import field_array, ptr, stream_handle
from jaxphys_amd.exceptions import ConfigurationError, NativeLibraryError, SimulationError

HERE = os.path.dirname(os.path.abspath(__file__))

EPI_GELU, EPI_LINEAR_F32, EPI_GELU_BWD, EPI_LINEAR_BF16 = 0, 1, 2, 3


class ConvDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("cin_pad", ctypes.c_int32), ("cout", ctypes.c_int32),
                ("cout_pad", ctypes.c_int32), ("y_cpad", ctypes.c_int32)]


EXPORTS = {
    "conv3d_fwd": (ctypes.c_int, [ctypes.POINTER(ConvDesc), ctypes.c_void_p, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                  ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_wgrad_work_floats": (ctypes.c_int64, [ctypes.POINTER(ConvDesc)]),
    "conv3d_wgrad": (ctypes.c_int, [ctypes.POINTER(ConvDesc), ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_flip_weights": (ctypes.c_int, [ctypes.POINTER(ConvDesc), ctypes.c_void_p,
                                           ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_pack_fields": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_apply_correction": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                               ctypes.c_void_p, ctypes.c_float, ctypes.c_int32,
                                               ctypes.c_void_p]),
    "conv3d_stencil_f32": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                          ctypes.c_void_p]),
    "conv3d_pack_fields_box": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_float, ctypes.c_float, ctypes.c_int32]
                               + [ctypes.c_int32] * 6 + [ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_stencil_f32_set": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                              ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_apply_closure": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                            ctypes.c_float, ctypes.c_void_p, ctypes.c_float,
                                            ctypes.c_int32, ctypes.c_void_p]),
    "conv3d_patch_add": (ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_void_p, ctypes.c_void_p]
                         + [ctypes.c_int32] * 6 + [ctypes.c_float, ctypes.c_void_p]),
    "conv3d_patch_restrict_h": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int32] * 7
                                + [ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_patch_correct": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
                             + [ctypes.c_int32] * 3 + [ctypes.c_void_p]
                             + [ctypes.c_int32] * 7 + [ctypes.c_void_p, ctypes.c_void_p]),
    "conv3d_last_error": (ctypes.c_char_p, []),
}

_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "libconv3d_hip.so")
        if not os.path.exists(path):
            raise NativeLibraryError(f"{path} is not built; run `python -m jaxphys_amd.build`")
        h = ctypes.CDLL(path)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = h
    return _LIB


def check(rc: int, what: str) -> None:
    if rc == 0:
        return
    msg = lib().conv3d_last_error().decode(errors="replace")
    if rc == -1:
        raise ConfigurationError(f"{what}: {msg}")
    raise SimulationError(f"{what}: {msg}")


def kpad_of(cin_pad: int) -> int:
    return ((27 * cin_pad + 31) // 32) * 32


def desc(shape_xyz, cin_pad, cout, cout_pad, y_cpad) -> ConvDesc:
    nx, ny, nz = shape_xyz
    return ConvDesc(nx, ny, nz, cin_pad, cout, cout_pad, y_cpad)


def fwd(d: ConvDesc, x, w, bias, epi, y, z_out=None, aux=None, stream=None) -> None:
    check(lib().conv3d_fwd(ctypes.byref(d), ptr(x), ptr(w), ptr(bias) or None, int(epi),
                           ptr(z_out) or None, ptr(aux) or None, ptr(y), stream_handle(stream)),
          "conv3d_fwd")


def wgrad_work_floats(d: ConvDesc) -> int:
    return int(lib().conv3d_wgrad_work_floats(ctypes.byref(d)))


def wgrad(d: ConvDesc, x, dy, dy_cpad, work, dw, db=None, stream=None) -> None:
    check(lib().conv3d_wgrad(ctypes.byref(d), ptr(x), ptr(dy), int(dy_cpad), ptr(work), ptr(dw),
                             ptr(db) or None, stream_handle(stream)), "conv3d_wgrad")


def flip_weights(d: ConvDesc, w, dy_cpad, wt, stream=None) -> None:
    check(lib().conv3d_flip_weights(ctypes.byref(d), ptr(w), int(dy_cpad), ptr(wt),
                                    stream_handle(stream)), "conv3d_flip_weights")


PACK_ASINH = 1


def pack_fields(shape_xyz, fields, eps, scale, eps_shift, eps_scale, out, flags=0,
                stream=None) -> None:
    nx, ny, nz = shape_xyz
    check(lib().conv3d_pack_fields(nx, ny, nz, field_array(fields), ptr(eps), ptr(scale),
                                   float(eps_shift), float(eps_scale), int(flags), ptr(out),
                                   stream_handle(stream)), "conv3d_pack_fields")


def pack_fields_box(shape_xyz, fields, eps, scale, eps_shift, eps_scale, box, out, flags=0,
                    stream=None) -> None:
    """``conv3d_pack_fields_box``: box = (x0, y0, z0, bnx, bny, bnz)."""
    nx, ny, nz = shape_xyz
    x0, y0, z0, bx, by, bz = (int(v) for v in box)
    check(lib().conv3d_pack_fields_box(nx, ny, nz, field_array(fields), ptr(eps), ptr(scale),
                                       float(eps_shift), float(eps_scale), int(flags), x0, y0,
                                       z0, bx, by, bz, ptr(out), stream_handle(stream)),
          "conv3d_pack_fields_box")


def apply_closure(shape_xyz, fields, s, sstride, yb, y0b, ybstride, box, mask, mask_all,
                  net_gain, scale, gain=1.0, margin=0, stream=None) -> None:
    """``conv3d_apply_closure`` (s, yb/y0b, mask may be None; box = (x0,
    y0, z0, bnx, bny, bnz) of yb)."""
    nx, ny, nz = shape_xyz
    b = (ctypes.c_int32 * 6)(*(int(v) for v in (box if box is not None else (0,) * 6)))
    check(lib().conv3d_apply_closure(nx, ny, nz, field_array(fields), ptr(s) or None,
                                     int(sstride), ptr(yb) or None, ptr(y0b) or None,
                                     int(ybstride), b if yb is not None else None,
                                     ptr(mask) or None, int(bool(mask_all)), float(net_gain),
                                     ptr(scale), float(gain), int(margin),
                                     stream_handle(stream)),
          "conv3d_apply_closure")


def patch_add(shape_xyz, fields, corr, origin_xyz, box_xyz, gain=1.0, stream=None) -> None:
    """``conv3d_patch_add``: fields += gain * corr (6, bz, by, bx) over the
    box at ``origin_xyz``."""
    nx, ny, nz = shape_xyz
    x0, y0, z0 = origin_xyz
    bx, by, bz = box_xyz
    check(lib().conv3d_patch_add(nx, ny, nz, field_array(fields), ptr(corr), int(x0), int(y0),
                                 int(z0), int(bx), int(by), int(bz), float(gain),
                                 stream_handle(stream)), "conv3d_patch_add")


def apply_correction(shape_xyz, fields, y, ystride, scale, gain=1.0, margin=0,
                     stream=None) -> None:
    nx, ny, nz = shape_xyz
    check(lib().conv3d_apply_correction(nx, ny, nz, field_array(fields), ptr(y), int(ystride),
                                        ptr(scale), float(gain), int(margin),
                                        stream_handle(stream)),
          "conv3d_apply_correction")


def stencil_f32(shape_xyz, fields, eps, scale, w, y, ystride, stream=None) -> None:
    """``conv3d_stencil_f32``: y[..][co] += the linear far-field operator of
    the six fields (w: 27*6*2*6 float32, [tap][field][plain, (eps-1)][co])."""
    nx, ny, nz = shape_xyz
    check(lib().conv3d_stencil_f32(nx, ny, nz, field_array(fields), ptr(eps), ptr(scale), ptr(w),
                                   ptr(y), int(ystride), stream_handle(stream)),
          "conv3d_stencil_f32")


def stencil_f32_set(shape_xyz, fields, eps, scale, w, y, ystride, box=None,
                    stream=None) -> None:
    """``conv3d_stencil_f32_set``: y[..][co] = the far-field operator's value
    on the cells of ``box`` (x0, y0, z0, bnx, bny, bnz; None: all)."""
    nx, ny, nz = shape_xyz
    b = (ctypes.c_int32 * 6)(*(int(v) for v in box)) if box is not None else None
    check(lib().conv3d_stencil_f32_set(nx, ny, nz, field_array(fields), ptr(eps), ptr(scale),
                                       ptr(w), ptr(y), int(ystride), b, stream_handle(stream)),
          "conv3d_stencil_f32_set")


def patch_restrict_h(fine_h, fine_shape_xyz, origin_xyz, side, out, stream=None) -> None:
    """``conv3d_patch_restrict_h``: out (3, side, side, side) = the fine H
    fields restricted to the coarse box at ``origin_xyz``."""
    nx, ny, nz = fine_shape_xyz
    x0, y0, z0 = origin_xyz
    check(lib().conv3d_patch_restrict_h(field_array(fine_h), nx, ny, nz, x0, y0, z0, int(side),
                                        ptr(out), stream_handle(stream)),
          "conv3d_patch_restrict_h")


def patch_correct(fine_e, fine_h, hprev, fine_shape_xyz, coarse, coarse_shape_xyz, origin_xyz,
                  side, corr, stream=None) -> None:
    """``conv3d_patch_correct``: corr (6, side, side, side) = restricted fine
    state minus the coarse fields in the box, and the coarse fields += corr."""
    nx, ny, nz = fine_shape_xyz
    cx, cy, cz = coarse_shape_xyz
    x0, y0, z0 = origin_xyz
    check(lib().conv3d_patch_correct(field_array(fine_e), field_array(fine_h), ptr(hprev), nx, ny,
                                     nz, field_array(coarse), cx, cy, cz, x0, y0, z0, int(side),
                                     ptr(corr), stream_handle(stream)),
          "conv3d_patch_correct")
