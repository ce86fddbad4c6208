"""Host <-> device copies of the C-ABI at sizes that use the pinned bounce
buffers (128 MiB and more): pf_put_state / pf_get_state round trip (exact in
the FP64 faithful form, whose state is the real fields themselves) and the
f_trans / inv_trans pair on one 256^3 field (the reference's normalisation
makes the pair the identity)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_state_round_trip_through_bounce_buffers(pf):
    nx, ny, nz, var = 256, 128, 128, 16        # phi: 537 MB, con: 34 MB
    rng = np.random.default_rng(7)
    phi = rng.standard_normal((nz, ny, nx, var))
    con = rng.standard_normal((nz, ny, nx))
    with pf.Simulation(pf.make_params(), nx, ny, nz, var, "fp64_faithful") as sim:
        sim.put_state(phi, con)
        gphi, gcon = sim.get_state()
    assert np.array_equal(gphi, phi)
    assert np.array_equal(gcon, con)


def test_f_trans_inv_trans_256_cubed(pf):
    n = 256
    rng = np.random.default_rng(3)
    x = rng.standard_normal((n, n, n))
    h = pf.f_trans(x)
    # spot check against numpy: one z-line of the spectrum
    ref = np.fft.rfftn(x) / float(np.float32(n ** 3))     # (z, y, kx)
    assert np.allclose(h[5, 7, :], ref[:, 7, 5], rtol=0, atol=1e-15)
    y = pf.inv_trans(h, n)
    assert np.abs(y - x).max() <= 1e-12
