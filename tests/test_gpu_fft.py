"""GPU transform parity: libpf's HIP R2C/C2R against numpy.fft and the oracle.

The reference transform is P3DFFT 2.7.9 over FFTW 3.3.10 (not available here):
forward e^{-i}, normalised by 1/(Nx*Ny*Nz) in f_trans (src/f_trans.f90:11-12),
inverse unnormalised (src/inv_trans.f90:10).  Tolerances: FP64 1e-13 relative to
the spectrum's max magnitude, FP32 2e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(8, 8, 8), (16, 12, 8), (32, 24, 16), (48, 16, 32), (64, 64, 64), (96, 32, 24),
          (160, 8, 16), (192, 12, 8), (256, 16, 8), (512, 8, 8), (8, 512, 8), (8, 8, 512),
          (128, 160, 96)]


def _torch():
    import torch
    assert torch.cuda.is_available(), "GPU test needs cuda:0"
    return torch


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("layout", ["xfast", "zfast"])
def test_r2c_c2r_vs_numpy(pf, shape, prec, layout):
    torch = _torch()
    nx, ny, nz = shape
    batch = 2
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    x = rng.standard_normal((batch, nz, ny, nx))
    ref = np.fft.rfftn(x, axes=(1, 2, 3)) / (nx * ny * nz)         # (b, nz, ny, nx/2+1)
    dt = torch.float64 if prec == "fp64" else torch.float32
    p = pf.PREC_FP64 if prec == "fp64" else pf.PREC_FP32
    lay = pf.LAYOUT_XFAST if layout == "xfast" else pf.LAYOUT_ZFAST
    mx = nx // 2 + 1
    dx = torch.tensor(x, dtype=dt, device="cuda")
    dspec = torch.empty((batch, nz * ny * mx * 2), dtype=dt, device="cuda")
    torch.cuda.synchronize()
    pf.fft_r2c_device(nx, ny, nz, batch, p, dx.data_ptr(), dspec.data_ptr(), lay, 1.0 / (nx * ny * nz))
    torch.cuda.synchronize()
    s = dspec.cpu().double().numpy().reshape(batch, -1, 2)
    s = s[..., 0] + 1j * s[..., 1]
    if layout == "xfast":
        got = s.reshape(batch, nz, ny, mx)
    else:
        got = s.reshape(batch, mx, ny, nz).transpose(0, 3, 2, 1)
    tol = 1e-13 if prec == "fp64" else 2e-6
    scale = np.abs(ref).max()
    assert np.abs(got - ref).max() / scale < tol
    # inverse (unnormalised): recovers x
    dback = torch.empty((batch, nz * ny * nx), dtype=dt, device="cuda")
    pf.fft_c2r_device(nx, ny, nz, batch, p, dspec.data_ptr(), dback.data_ptr(), lay)
    torch.cuda.synchronize()
    back = dback.cpu().double().numpy().reshape(batch, nz, ny, nx)
    tol2 = 1e-13 if prec == "fp64" else 5e-6
    assert np.abs(back - x).max() / np.abs(x).max() < tol2


def test_c2r_ignores_imag_of_dc_and_nyquist_planes(pf):
    """A C2R input whose kx=0 / kx=nx/2 entries carry a non-Hermitian part: the
    HIP path must agree with the oracle (Hermitian projection of each x line)."""
    import oracle
    nx, ny, nz = 16, 8, 12
    rng = np.random.default_rng(7)
    h = rng.standard_normal((nz, ny, nx // 2 + 1)) + 1j * rng.standard_normal((nz, ny, nx // 2 + 1))
    ref = oracle.fft_c2r(h, nx)
    got = pf.inv_trans(np.ascontiguousarray(h.transpose(2, 1, 0)), nx)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-13


@pytest.mark.parametrize("shape", [(16, 16, 16), (24, 32, 8), (64, 48, 96)])
def test_f_trans_inv_trans_host_dropins(pf, oracle, shape):
    nx, ny, nz = shape
    rng = np.random.default_rng(3)
    x = rng.standard_normal((nz, ny, nx))
    got = pf.f_trans(x)                                  # (mx, ny, nz) == Fortran (nz, ny, mx)
    ref = oracle.fft_r2c(x)                              # (nz, ny, mx)
    assert np.abs(got.transpose(2, 1, 0) - ref).max() / np.abs(ref).max() < 1e-13
    back = pf.inv_trans(got, nx)
    assert np.abs(back - x).max() < 1e-12
