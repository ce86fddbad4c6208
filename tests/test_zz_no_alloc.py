"""pf_fft_r2c_noalloc: the forward transform with its scratch in a static device
array instead of an allocation.  Called through the C-ABI (ctypes) and checked
against pf_fft_r2c (the entry point it varies) on the same inputs: the two share
their kernels and twiddles, so the spectra must be bit-identical.  Two calls in
a row, the second on a smaller grid, so the scratch left by the first call must
not leak into the second; a request above the scratch limit is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test needs cuda:0"
    return torch


def _noalloc(pf):
    L = pf.lib()
    f = L.pf_fft_r2c_noalloc
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                  C.c_double, C.c_void_p]
    f.restype = C.c_int
    return f


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_noalloc_r2c_equals_r2c_twice(pf, prec):
    torch = _torch()
    fn = _noalloc(pf)
    dt = torch.float64 if prec == "fp64" else torch.float32
    p = pf.PREC_FP64 if prec == "fp64" else pf.PREC_FP32
    # larger first, then smaller: the second call reads none of the first's scratch
    cases = [((64, 48, 32), 3, pf.LAYOUT_ZFAST), ((16, 24, 8), 2, pf.LAYOUT_XFAST)]
    for (nx, ny, nz), batch, lay in cases:
        print(f"noalloc r2c {prec} {nx}x{ny}x{nz} batch {batch} layout {lay}", flush=True)
        rng = np.random.default_rng(nx + ny + nz + batch)
        x = torch.tensor(rng.standard_normal((batch, nz, ny, nx)), dtype=dt, device="cuda")
        mx = nx // 2 + 1
        a = torch.empty((batch, nz * ny * mx * 2), dtype=dt, device="cuda")
        b = torch.full_like(a, float("nan"))
        scale = 1.0 / (nx * ny * nz)
        torch.cuda.synchronize()
        pf.fft_r2c_device(nx, ny, nz, batch, p, x.data_ptr(), a.data_ptr(), lay, scale)
        torch.cuda.synchronize()
        rc = fn(nx, ny, nz, batch, p, x.data_ptr(), b.data_ptr(), lay, scale, None)
        assert rc == 0, pf.lib().pf_last_error().decode()
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        # and the spectrum is the transform (numpy, 1e-13 / 2e-6 of the max)
        ref = np.fft.rfftn(x.cpu().double().numpy(), axes=(1, 2, 3)) * scale
        s = b.cpu().double().numpy().reshape(batch, -1, 2)
        s = s[..., 0] + 1j * s[..., 1]
        got = s.reshape(batch, nz, ny, mx) if lay == pf.LAYOUT_XFAST else \
            s.reshape(batch, mx, ny, nz).transpose(0, 3, 2, 1)
        assert np.abs(got - ref).max() / np.abs(ref).max() < (1e-13 if prec == "fp64" else 2e-6)


def test_noalloc_r2c_refuses_above_its_scratch(pf):
    torch = _torch()
    fn = _noalloc(pf)
    x = torch.empty(16, dtype=torch.float64, device="cuda")  # never read: the size check comes first
    rc = fn(256, 256, 256, 1, pf.PREC_FP64, x.data_ptr(), x.data_ptr(), pf.LAYOUT_XFAST, 1.0, None)
    assert rc == pf.PF_EINVAL
    assert "scratch" in pf.lib().pf_last_error().decode()
