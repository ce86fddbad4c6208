"""GPU modes against closed-form solutions of the scheme (tests/closed_forms.py):
a single Cahn-Hilliard mode with and without the variable-mobility
correction, Allen-Cahn harmonics, and the gradient energy of a single mode.
These pin the GPU path without going through the oracle.

Tolerances: FP64 forms 1e-13 absolute on fields of size O(1); FP32-mixed 2e-6."""
import numpy as np
import pytest

from closed_forms import ac_harmonics_case, ch_mode_case, single_mode_fint_case

pytestmark = pytest.mark.gpu
MODES = ["fp64_faithful", "fp64_reduced", "fp32_mixed"]
TOL = {"fp64_faithful": 1e-13, "fp64_reduced": 1e-13, "fp32_mixed": 2e-6}


def _run(pf, p, phi, con, mode, nsteps):
    nz, ny, nx, var = phi.shape
    with pf.Simulation(p, nx, ny, nz, var, mode) as sim:
        sim.put_state(phi, con)
        dg = sim.step(nsteps, diag_every=1)
        return sim.get_state() + (dg,)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("wave,dgb", [((1, 2, 1), 0.0), ((2, 1, 3), 0.3), ((0, 3, 1), 0.5)])
def test_ch_single_mode(pf, mode, wave, dgb):
    p = pf.make_params(t_step=0.01, grad_coeff_c=0.7, D_L=1.3, D_gb=dgb, ap=2.0, con_0_ppt=0.3,
                       mob_phi=0.8)
    phi, con, ephi, econ, _ = ch_mode_case(p, wave, (16, 12, 8), 6)
    gphi, gcon, _ = _run(pf, p, phi, con, mode, 6)
    print(mode, wave, dgb, np.abs(gcon - econ).max(), np.abs(gphi - ephi).max())
    assert np.abs(gcon - econ).max() < TOL[mode]
    assert np.abs(gphi - ephi).max() < TOL[mode]


@pytest.mark.parametrize("mode", MODES)
def test_allen_cahn_harmonics(pf, mode):
    p = pf.make_params(t_step=0.02, grad_coeff_phi=0.9, mob_phi=1.2)
    phi, con, expect = ac_harmonics_case(p, (2, 1, 0), (24, 8, 8))
    gphi, _, _ = _run(pf, p, phi, con, mode, 1)
    assert np.abs(gphi - expect).max() < TOL[mode]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("wave", [(1, 0, 0), (2, 3, 1), (8, 0, 0)])
def test_gradient_energy_single_mode(pf, mode, wave):
    p = pf.make_params(grad_coeff_phi=0.8, grad_coeff_c=0.6)
    phi, con, expect = single_mode_fint_case(p, wave, (16, 12, 8))
    _, _, dg = _run(pf, p, phi, con, mode, 1)
    tol = 1e-12 if mode != "fp32_mixed" else 1e-5
    assert dg[0]["f_int"] == pytest.approx(expect, rel=tol, abs=1e-20 if expect == 0 else 0)
