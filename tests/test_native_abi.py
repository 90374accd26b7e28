"""The C ABI library loads and exports every entry point include/*.h
declares; host-side validation works without a GPU (no compute calls)."""

import ctypes
import os
import re

import pytest

from jaxphys_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return set(re.findall(r"\b([a-z][a-z0-9_]*)\s*\(", text)) - {"if", "sizeof"}


def test_yee_library_exports_all_declared_symbols():
    lib = _native.lib()
    names = _declared("yee.h")
    assert "yee_run" in names and "yee_plan_create" in names
    for n in names:
        assert hasattr(lib, n), n
    assert set(_native.EXPORTS) == names


def _desc(**kw):
    d = _native.YeeGridDesc()
    d.nx, d.ny, d.nz = kw.get("n", (16, 12, 10))
    d.boundary = kw.get("boundary", _native.BOUNDARY_WALLS)
    d.ch = 0.1
    d.cb = kw.get("cb", 0x1000)  # never dereferenced by plan creation
    d.npml = kw.get("npml", 0)
    d.cpml = kw.get("cpml", None)
    d.sponge_sigma = kw.get("sponge", None)
    d.mat = kw.get("mat", None)
    d.palette = kw.get("palette", None)
    return d


def test_psi_floats_formula():
    d = _desc(n=(20, 18, 16), boundary=_native.BOUNDARY_CPML, npml=4, cpml=0x2000)
    n = _native.lib().yee_psi_floats(ctypes.byref(d))
    tp = 8
    assert n == 4 * (16 * tp * 20 + tp * 18 * 20 + 16 * 18 * tp)
    assert _native.lib().yee_psi_floats(ctypes.byref(_desc())) == 0


@pytest.mark.parametrize("kw,msg", [
    (dict(n=(1, 12, 10)), "at least 2"),
    (dict(boundary=7), "unknown boundary"),
    (dict(cb=None), "cb"),
    (dict(boundary=_native.BOUNDARY_SPONGE), "sponge"),
    (dict(boundary=_native.BOUNDARY_CPML, npml=6, cpml=0x2000), "overlap"),
    (dict(boundary=_native.BOUNDARY_CPML, npml=2), "tables"),
    (dict(mat=0x3000), "palette"),
])
def test_plan_create_rejects_bad_descriptors(kw, msg):
    d = _desc(**kw)
    h = ctypes.c_void_p()
    rc = _native.lib().yee_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == _native.YEE_EINVAL
    assert msg in _native.lib().yee_last_error().decode()
    with pytest.raises(Exception):
        _native.check(rc, "yee_plan_create")


def test_plan_and_source_validation_host_only():
    from jaxphys_amd import ConfigurationError

    plan = _native.Plan(_desc())
    import numpy as np
    with pytest.raises(ConfigurationError, match="out of grid"):
        plan.set_sources([2], [(16, 0, 0)], np.zeros((3, 1), np.float32), 3)
    with pytest.raises(ConfigurationError, match="component"):
        plan.set_sources([4], [(1, 1, 1)], np.zeros((3, 1), np.float32), 3)
    # (a valid call uploads the table to the device: exercised by the GPU tests)


def test_rejected_set_sources_leaves_the_plan_usable_host_only():
    """A rejected yee_set_sources changes nothing: repeated rejected calls
    (bad position after good ones in the same call, bad component, null
    arrays) each return EINVAL with their message, and the plan is then
    destroyed cleanly.  The GPU tests check that a run after a rejected call
    uses the previous sources."""
    import numpy as np

    lib = _native.lib()
    plan = _native.Plan(_desc())
    comp = np.array([2, 2], np.int32)
    pos = np.array([1, 1, 1, 99, 0, 0], np.int32)
    tab = np.zeros(6, np.float32)
    for _ in range(2):
        rc = lib.yee_set_sources(plan.handle, 2, comp.ctypes.data, pos.ctypes.data,
                                 tab.ctypes.data, 3)
        assert rc == _native.YEE_EINVAL and b"out of grid" in lib.yee_last_error()
    bad = np.array([5], np.int32)
    rc = lib.yee_set_sources(plan.handle, 1, bad.ctypes.data, pos.ctypes.data,
                             tab.ctypes.data, 3)
    assert rc == _native.YEE_EINVAL and b"component" in lib.yee_last_error()
    rc = lib.yee_set_sources(plan.handle, 1, None, None, None, 3)
    assert rc == _native.YEE_EINVAL and b"null" in lib.yee_last_error()
    plan.close()
    plan.close()


def test_tm2d_library_exports_all_declared_symbols():
    from jaxphys_amd import _tm2d

    lib = _tm2d.lib()
    names = _declared("tm2d.h")
    assert {"tm2d_step", "tm2d_run", "tm2d_last_error"} <= names
    for n in names:
        assert hasattr(lib, n), n
    assert set(_tm2d.EXPORTS) == names


@pytest.mark.parametrize("kw,msg", [
    (dict(nx=2), "at least 3x3"),
    (dict(boundary=5), "unknown boundary"),
    (dict(npml=3), "damping table"),
    (dict(n_src=1), "src_rows"),
    (dict(nx=70000, ny=70000), "2^31"),
])
def test_tm2d_rejects_bad_descriptors_host_only(kw, msg):
    from jaxphys_amd import _tm2d

    d = _tm2d.Tm2dDesc()
    d.nx, d.ny = kw.get("nx", 16), kw.get("ny", 12)
    d.boundary = kw.get("boundary", _tm2d.ABSORBING)
    d.npml = kw.get("npml", 0)
    d.n_src = kw.get("n_src", 0)
    f = _tm2d.fields3([None, None, None])
    rc = _tm2d.lib().tm2d_run(ctypes.byref(d), f, f, 0, 1, None)
    assert rc == _tm2d.TM2D_EINVAL
    assert msg in _tm2d.lib().tm2d_last_error().decode()


def test_tm2d_rejects_aliased_buffers_host_only():
    from jaxphys_amd import _tm2d

    d = _tm2d.Tm2dDesc()
    d.nx, d.ny, d.boundary = 16, 12, _tm2d.REFLECTING
    a = (ctypes.c_void_p * 3)(0x1000, 0x2000, 0x3000)
    b = (ctypes.c_void_p * 3)(0x4000, 0x2000, 0x5000)  # shares one buffer with a
    rc = _tm2d.lib().tm2d_run(ctypes.byref(d), a, b, 0, 2, None)
    assert rc == _tm2d.TM2D_EINVAL and "differ" in _tm2d.lib().tm2d_last_error().decode()


def test_conv3d_library_exports_all_declared_symbols():
    from jaxphys_amd import _conv

    lib = _conv.lib()
    names = _declared("conv3d.h")
    assert {"conv3d_fwd", "conv3d_wgrad", "conv3d_pack_fields"} <= names
    for n in names:
        assert hasattr(lib, n), n
    assert set(_conv.EXPORTS) == names


def test_conv3d_rejects_bad_channel_configs_host_only():
    from jaxphys_amd import _conv

    d = _conv.desc((16, 8, 4), 12, 32, 32, 32)   # 12 input channels: not a multiple of 8
    rc = _conv.lib().conv3d_fwd(ctypes.byref(d), 0x1000, 0x2000, None, 0, None, None, 0x3000,
                                None)
    assert rc == -1 and "channel" in _conv.lib().conv3d_last_error().decode()
    rc = _conv.lib().conv3d_apply_correction(4, 4, 4, (ctypes.c_void_p * 6)(), 0x1000, 8,
                                             0x2000, 1.0, -1, None)
    assert rc == -1
    rc = _conv.lib().conv3d_stencil_f32(4, 4, 4, (ctypes.c_void_p * 6)(), 0x1000, 0x2000,
                                        0x3000, 0x4000, 8, None)
    assert rc == -1 and "null field" in _conv.lib().conv3d_last_error().decode()
    rc = _conv.lib().conv3d_stencil_f32(4, 4, 4, (ctypes.c_void_p * 6)(*[0x1000] * 6), 0x1000,
                                        0x2000, 0x3000, 0x4000, 5, None)
    assert rc == -1


def test_patch_kernels_reject_bad_boxes():
    """The near-source patch entry points check their arguments before any
    GPU work: an odd fine grid or a box outside the grid is -1 with a
    message (pointer values are never dereferenced on these paths)."""
    import ctypes

    from jaxphys_amd import _conv

    lib = _conv.lib()
    fake = (ctypes.c_void_p * 6)(*([16] * 6))
    out = ctypes.c_void_p(16)
    assert lib.conv3d_patch_restrict_h(fake, 33, 32, 32, 0, 0, 0, 4, out, None) == -1
    assert b"even" in lib.conv3d_last_error()
    assert lib.conv3d_patch_restrict_h(fake, 32, 32, 32, 14, 0, 0, 4, out, None) == -1
    assert b"outside" in lib.conv3d_last_error()
    assert lib.conv3d_patch_correct(fake, fake, out, 32, 32, 32, fake, 16, 16, 8, 0, 0, 6, 4,
                                    out, None) == -1
    assert b"outside" in lib.conv3d_last_error()
