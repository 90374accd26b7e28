"""The allocation-free energy entry point (yee_energy_fp64_noalloc) against
yee_energy_fp64: the same kernels and summation order, so the same bits, on
two calls in a row (the second on a smaller grid, so stale partial sums of
the first call would show), called through the C ABI."""

import pytest
import torch

from jaxphys_amd import _native


def _fields(shape_xyz, seed, dev):
    nx, ny, nz = shape_xyz
    g = torch.Generator(device="cpu").manual_seed(seed)
    return [torch.randn((nz, ny, nx), generator=g).to(dev) for _ in range(6)]


@pytest.mark.gpu
def test_noalloc_energy_equals_the_workspace_version(gpu):
    out = torch.full((2,), float("nan"), dtype=torch.float64, device=gpu)
    for shape, seed in (((96, 80, 72), 0), ((20, 9, 13), 1)):
        f = _fields(shape, seed, gpu)
        eps = 1.0 + 3.0 * torch.rand(f[0].shape, generator=torch.Generator().manual_seed(seed + 7)
                                     ).to(gpu)
        for e in (eps, None):
            ref = _native.energy_fp64(shape, f, e).clone()
            _native.energy_fp64_noalloc(shape, f, e, out)
            torch.cuda.synchronize()
            assert torch.equal(out, ref), (shape, e is None, out.tolist(), ref.tolist())
            # and against float64 sums in torch (order differs: ~1e-15)
            we = sum(((e.double() if e is not None else 1.0) * a.double() ** 2).sum()
                     for a in f[:3])
            wh = sum((a.double() ** 2).sum() for a in f[3:])
            assert torch.allclose(out, torch.stack([we, wh]), rtol=1e-12, atol=0)


@pytest.mark.gpu
def test_noalloc_energy_rejects_bad_arguments(gpu):
    lib = _native.lib()
    f = _fields((8, 8, 8), 2, gpu)
    out = torch.zeros(2, dtype=torch.float64, device=gpu)
    big = 1 << 13  # 2^39 cells > YEE_NOALLOC_MAX_CELLS
    rc = lib.yee_energy_fp64_noalloc(big, big, big, _native.field_array(f), None,
                                     _native.ptr(out), None)
    assert rc == _native.YEE_EINVAL and b"maximum" in lib.yee_last_error()
    rc = lib.yee_energy_fp64_noalloc(8, 8, 8, _native.field_array(f), None, None, None)
    assert rc == _native.YEE_EINVAL
    assert torch.equal(out, torch.zeros_like(out))
