"""config.yaml -> input.txt plumbing (tests/scaling_tests.py:53-69 and
tests/run_phase_sim.py:112-128 of the reference) and the ensemble's parameter
draws.  CPU only."""
import os

import pytest

from driver_util import DRIVER, run_driver

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def rc(pf):
    import pf_amd.runconfig as rc
    return rc


def test_config_list_applies_swaps(rc):
    cfg = rc.load_config(os.path.join(GOLD, "config_sample.yaml"))
    lst = rc.config_list(cfg, "size_test")
    assert [(c["Nx"], c["Ny"], c["Nz"]) for c in lst] == [(16, 16, 16), (24, 32, 16)]
    assert lst[0]["d_gb"] == 0.5 and lst[1]["d_gb"] == 0.0
    assert cfg["baseline_config"]["Nx"] == 32  # the baseline is not modified


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="var_diff_hip.x not built")
def test_written_input_is_read_by_the_driver(rc, tmp_path):
    cfg = rc.load_config(os.path.join(GOLD, "config_sample.yaml"))
    c = rc.config_list(cfg, "size_test")[1]
    path = rc.write_input(str(tmp_path), c)
    txt = open(path).read().splitlines()
    assert txt[0].startswith("1 0 ") and txt[2].startswith("48 70 48 -1") and txt[-1].startswith("'circle'")
    out = tmp_path / "out"
    r = run_driver(str(tmp_path), path, str(out), env={"PF_DRY_RUN": "1"})
    assert r.returncode == 0, r.stderr
    # gamma_matrix.dat reflects sigma_1/sigma_2 for var = 3
    rows = [l.split() for l in open(out / "gamma_matrix.dat") if l.strip()]
    assert len(rows) == 3 and float(rows[0][2]) == 1.5


def test_ensemble_draws(rc, pf):
    ms = rc.ensemble_members(pf, n=12)
    assert [m["grid"] for m in ms[:6]] == [64, 96, 128, 160, 192, 64]
    r = pf.Ran2(-1)
    for m in ms:
        k, mo, lo = 0.25 + 1.75 * r(), 0.5 + 1.5 * r(), 0.5 + 1.5 * r()
        assert (m["kappa"], m["M"], m["L"]) == (k, mo, lo)
        assert 0.25 <= k <= 2.0 and 0.5 <= mo <= 2.0 and 0.5 <= lo <= 2.0
        assert m["params"].grad_coeff_c == k and m["params"].D_L == mo and m["params"].mob_phi == lo
